"""The positivity fixer of adv_run (adv_mod.F90:236-258) walks a row from
west to east and fixes every value still negative when reached.  The gfx950
kernel (k_adv_fix in gmcore_amd/csrc/gmc_adv.hip) instead splits the
originally negative columns into clusters whose members are at most two
columns apart and walks the clusters in parallel, each west to east.  This
test checks that claim on the CPU: for rows dense with negative values
(adjacent ones, rows that start or end negative), walking the clusters in any
order gives the bits and the fix count of the sequential walk."""
import numpy as np


def _fix_at(q, m, i, seam):
    n = len(q)
    im, ip = (n - 1 if i == 0 else i - 1), (0 if i == n - 1 else i + 1)
    qm0 = q[i] * m[i]
    qm1 = (seam[0] if i == 0 else q[im]) * m[im]
    qm2 = (seam[1] if i == n - 1 else q[ip]) * m[ip]
    h = 0.5 * qm0
    if qm1 >= h and qm2 >= h:
        a, b = (qm1 - h) / m[im], (qm2 - h) / m[ip]
        if i == 0:
            seam[0] = a
        else:
            q[im] = a
        if i == n - 1:
            seam[1] = b
        else:
            q[ip] = b
    elif qm1 > qm0:
        a = (qm1 - qm0) / m[im]
        if i == 0:
            seam[0] = a
        else:
            q[im] = a
    elif qm2 > qm0:
        b = (qm2 - qm0) / m[ip]
        if i == n - 1:
            seam[1] = b
        else:
            q[ip] = b
    q[i] = 0.0


def sequential(q, m, west, east):
    q = q.copy()
    seam = [west, east]
    nfix = 0
    for i in range(len(q)):
        if q[i] < 0:
            _fix_at(q, m, i, seam)
            nfix += 1
    return q, nfix


def clustered(q, m, west, east, rng):
    q = q.copy()
    n = len(q)
    neg = q < 0
    starts = [i for i in range(n) if neg[i] and not (i >= 1 and neg[i - 1]) and not (i >= 2 and neg[i - 2])]
    rng.shuffle(starts)
    seam = [west, east]
    nfix = 0
    for i in starts:
        while True:
            if q[i] < 0:
                _fix_at(q, m, i, seam)
                nfix += 1
            if i + 1 < n and neg[i + 1]:
                i += 1
            elif i + 2 < n and neg[i + 2]:
                i += 2
            else:
                break
    return q, nfix


def test_clusters_match_sequential_walk():
    rng = np.random.default_rng(20)
    total = 0
    for trial in range(400):
        n = int(rng.integers(3, 80))
        frac = rng.uniform(0.05, 0.9)
        q = rng.uniform(0.0, 1.0, n)
        q[rng.random(n) < frac] *= -rng.uniform(0.01, 3.0)
        if trial % 4 == 0:
            q[0] = -abs(q[0]) - 0.1
        if trial % 4 == 1:
            q[-1] = -abs(q[-1]) - 0.1
        m = rng.uniform(0.5, 2.0, n)
        west, east = rng.uniform(-0.5, 1.0, 2)
        a, na = sequential(q, m, west, east)
        b, nb = clustered(q, m, west, east, rng)
        assert na == nb
        assert np.array_equal(a, b), trial
        total += na
    assert total > 1000
