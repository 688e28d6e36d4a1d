#!/usr/bin/env python3
"""Benchmark: FP64 shallow-water pc2 step at 0.25 degrees on one MI355X.

Workload (the headline configuration): 1440 x 721 C-grid (poles as cell
rows), FP64, polar filter on, pc2, dt = 60 s, seeded balanced initial state
(random spherical harmonics l <= 60, 40 m/s peak wind, mean geopotential
5.0e4 m2/s2, flat topography).  A step is one gmc_step: time_integrator
(three forward-backward stages) through libgmcore_hip.so, replayed from a
captured HIP graph, inputs resident in HBM.

Prints one JSON line.  value = grid-point-steps per second (nlon * nlat per
step).  roofline: the dominant kernel's algorithmic bytes per launch over its
mean launch time from gmc_kernel_times (a start and an end HIP event per
launch on the library's stream, stamped by the dispatch) against 8 TB/s, with the HBM traffic per launch measured by
rocprofv3 counters.  step_roofline: the algorithmic bytes of the whole step
(35 FP64 field passes per grid point, 280 B/gp) over the step time.  cpu_baseline: the oracle (a scalar C
restatement of the reference, gcc -O3, one thread as the reference runs) timed
on a bounded sample of the same workload; cpu_baseline_all_cores: the same
code with its row loops on OpenMP threads over the usable host CPUs.
targets: the two additions' target figures from the same run (the adjoint
gradients' cost over their forward windows, the batched ensemble's speedup
over its members run one at a time); skipped with --no-cpu or --no-targets.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_GP_STEP = 280.0          # 35 FP64 field passes per pc2 step
# compulsory bytes per grid point of one launch of each kernel class, mean
# over the three stages (k_tend reads 4 fields in stage 1 and 5 in stages
# 2-3 and writes 3; k_gzf reads 1, writes 2; k_uvf reads 5, writes 4)
KERNEL_BYTES_PER_GP = {"k_tend": 8.0 * (7 + 8 + 8) / 3, "k_gzf": 24.0, "k_uvf": 72.0}
PROFILES = os.path.join(ROOT, "profiles")


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` as measured by rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE, read from the `traffic-data:` line of the
    profile summaries under profiles/ (per-launch means in KiB and the factors
    a device copy of known size calibrates them with).  None when no summary
    holds the pair."""
    try:
        names = sorted(f for f in os.listdir(PROFILES) if f.endswith(".md"))
    except OSError:
        return None
    for name in names:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if not line.startswith("traffic-data:"):
                    continue
                kv = dict(t.split("=", 1) for t in line.split()[1:] if "=" in t)
                if kv.get("workload") == workload and kv.get("kernel") == kernel:
                    return 1024.0 * (float(kv["read_factor"]) * float(kv["fetch_kib"])
                                     + float(kv["write_factor"]) * float(kv["write_kib"]))
    return None

# FFSL tracer advection, 0.5 degrees, 8 tracers, dt = 900 s, one day
ADV_CONFIGS = {
    "ffsl05": dict(nlon=720, nlat=361, dt=900.0, ntracers=8, lmax=10, wind_seed=2, tracer_seed=3,
                   vmax=50.0, hw=104, steps=96),
}
# gradient of a final-state tracer misfit over a one-day FFSL window at the
# ffsl05 workload, against the same window's forward
ADV_ADJ_CONFIGS = {
    "adjffsl05": dict(base="ffsl05", window=96, ckpt=8, bg_seed=4),
}
# compulsory bytes per tracer-grid-point of an advection step: every tracer
# read and written once, the 4 wind fields (u, v, mfx, mfy) shared
ADV_BYTES_PER_TRACER_GP = lambda ntr: (2 * ntr + 4) * 8.0 / ntr

# batched ensemble: 64 members of uneven resolution, each with its stable dt
ENS_CONFIGS = {
    "ens64": dict(members=[(180, 91, 600.0, 32), (360, 181, 300.0, 24), (720, 361, 150.0, 8)],
                  lmax=20, perturb=0.01, steps=100),
}

# adjoint: 4D-Var-style gradient over a window at 1 degree
ADJ_CONFIGS = {
    "adj1": dict(nlon=360, nlat=181, dt=300.0, lmax=30, truth_seed=4, bg_perturb=(0.01, 6), nobs=5000,
                 obs_seed=5, sigma_o=50.0, sigma_b=1.0, window=48, ckpt=8),
}

# hydrostatic column operators: 1 degree, 32 levels of the equal-eta
# coordinate, a stratified column state
BARO_CONFIGS = {
    "baro1": dict(nlon=360, nlat=181, nlev=32, ptop=219.4, reps=20),
}
# algorithmic FP64 fields per column of each operator, as (full-level
# passes, half-level passes, single-level passes), reads and writes
BARO_FIELDS_PER_OP = {
    "calc_ph": (1, 2, 1), "calc_t": (3, 0, 0), "calc_gz_lev": (3, 1, 1), "calc_dphsdt": (2, 0, 1),
    "calc_we_lev": (2, 4, 1), "calc_wedudlev_wedvdlev": (6, 2, 0), "pgf_lin97": (2, 2, 0),
    "calc_ptf_lev": (2, 3, 0), "calc_ptf_hflx": (13, 0, 0), "small_filter2": (4, 0, 0),
}

CONFIGS = {
    "sw025": dict(nlon=1440, nlat=721, dt=60.0, lmax=60, seed=1),
    "sw2": dict(nlon=180, nlat=91, dt=600.0, lmax=20, seed=0),
    "sw1": dict(nlon=360, nlat=181, dt=300.0, lmax=30, seed=4),
    "sw05": dict(nlon=720, nlat=361, dt=150.0, lmax=40, seed=2),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_setup(cfg):
    from gmcore_amd.host import Setup
    from gmcore_amd.namelist import Namelist
    nl = Namelist(num_lon=cfg["nlon"], num_lat=cfg["nlat"], dt_dyn=cfg["dt"], time_scheme="pc2")
    return Setup(nl)


def make_ic(st, cfg):
    from gmcore_amd import init_cond
    nl = st.nl
    return init_cond.seeded_state(nl.num_lon, nl.num_lat, st.mesh.metrics, cfg["lmax"], cfg["seed"],
                                  omega=nl.omega, radius=nl.radius)


def host_cpu():
    """CPU model, logical CPUs of the machine and CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable, "mpi_ranks": 1}


def cpu_baseline(st, ic, cfg, budget_s=15.0, omp=False):
    """Oracle (C restatement of the reference, gcc -O3) on the same workload:
    steps until about budget_s seconds of CPU work, after one untimed step.
    omp=False: one thread, the reference's own execution model (one MPI rank,
    OpenMP off); omp=True: the same code with its row loops on OpenMP
    threads (OMP_NUM_THREADS, else every usable CPU), bit-identical."""
    import oracle
    info = host_cpu()
    if omp and "OMP_NUM_THREADS" not in os.environ:
        os.environ["OMP_NUM_THREADS"] = str(info["usable_cpus"])
    threads = int(os.environ.get("OMP_NUM_THREADS", info["usable_cpus"])) if omp else 1
    o = oracle.SWOracle(st.mesh_struct, st.cfg, omp=omp)
    for k, a in zip(("u_lon", "v_lat", "gz", "gzs"), ic):
        o.set(1, k, a)
    o.prepare(1)
    o.step(cfg["dt"], 1)
    n, t0 = 0, time.perf_counter()
    while True:
        o.step(cfg["dt"], 1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 400:
            break
    o.close()
    gp = cfg["nlon"] * cfg["nlat"]
    build = "liboracle_omp.so (gcc -O3 -fopenmp -ffp-contract=off)" if omp else \
        "liboracle.so (gcc -O3 -ffp-contract=off)"
    return {"value": gp * n / el, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} pc2 steps of the same {cfg['nlon']}x{cfg['nlat']} workload after 1 untimed "
                      f"step, {el:.1f} s, {build}, {threads} thread(s), 1 process; the mass batch's "
                      "accum_uv_cell, whose output the SW step never reads, is not run",
            **info}


def cpu_baseline_adv(st, qs, ub, vb, cfg, budget_s=15.0):
    """Oracle (scalar C, one core) on the same advection workload: steps
    until about budget_s seconds, after the accumulate-only first call."""
    import oracle
    from gmcore_amd import init_cond
    o = oracle.AdvOracle(st.mesh_struct, st.cfg)
    for l, q in enumerate(qs):
        o.set_tracer(1, l, q)
    nlat, nlon = cfg["nlat"], cfg["nlon"]
    ones = np.ones((nlat, nlon))
    sc = init_cond.wind_scales(200, cfg["dt"])
    n, el = 0, 0.0
    for step in range(200):
        u, v = ub * sc[step], vb * sc[step]
        o.set_wind(u, v, u, v, ones)
        t0 = time.perf_counter()
        o.run(step)
        if step > 0:
            el += time.perf_counter() - t0
            n += 1
        if el >= budget_s:
            break
    o.close()
    tgp = cfg["nlon"] * cfg["nlat"] * cfg["ntracers"]
    return {"value": tgp * n / el, "unit": "tracer-grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} advection steps of the same {nlon}x{nlat}x{cfg['ntracers']}-tracer workload "
                      f"after the accumulate-only first call, {el:.1f} s of adv_run, oracle built "
                      "gcc -O3 -ffp-contract=off, 1 thread"}


def run_adv(args):
    """FFSL advection benchmark: one step = adv_run on all tracers (set_uv of
    the prescribed flow included), device resident."""
    import ctypes as C
    from gmcore_amd import init_cond
    from gmcore_amd.adv import Advection, AdvSetup
    from gmcore_amd.namelist import Namelist
    cfg = ADV_CONFIGS[args.config]
    nlon, nlat, dt, ntr = cfg["nlon"], cfg["nlat"], cfg["dt"], cfg["ntracers"]
    nl = Namelist(num_lon=nlon, num_lat=nlat, dt_adv=dt, dt_dyn=dt, limiter_type="mono")
    st = AdvSetup(nl, ntr, lon_halo_width=cfg["hw"])
    m = st.mesh.metrics
    ub, vb = init_cond.deform_wind_basis(nlon, nlat, m, cfg["lmax"], cfg["wind_seed"], cfg["vmax"])
    qs = init_cond.tracer_set(nlon, nlat, m, cfg["tracer_seed"], ntr)
    g = Advection(st)
    for l, q in enumerate(qs):
        g.set_tracer(1, l, q)
    g.set_wind_basis(ub, vb)
    ac = st.mesh["area_cell"][:, None]
    import math
    m0 = [math.fsum((q * ac).ravel()) for q in qs]
    total = args.warmup + args.steps + 1
    sc = init_cond.wind_scales(total + 8, dt)
    g.drive(0, sc[:1 + args.warmup])
    g.synchronize()
    hip = C.CDLL("libamdhip64.so")
    stream = C.c_void_p()
    assert g.lib.gmc_adv_stream(g.h, C.byref(stream)) == 0
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(ev0)) == 0 and hip.hipEventCreate(C.byref(ev1)) == 0
    w0 = time.perf_counter()
    hip.hipEventRecord(ev0, stream)
    g.drive(1 + args.warmup, sc[1 + args.warmup:total])
    hip.hipEventRecord(ev1, stream)
    hip.hipEventSynchronize(ev1)
    wall = time.perf_counter() - w0
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    ms_step = ms.value / args.steps
    qend = [g.tracer(g.old, l) for l in range(ntr)]
    drift = max(abs(math.fsum((q * ac).ravel()) - m0[l]) / m0[l] for l, q in enumerate(qend[:-1]))
    qmin = min(float(q.min()) for q in qend)
    # per-kernel breakdown
    kt = np.zeros(8)
    nk = C.c_int32()
    from gmcore_amd import capi
    rc = g.lib.gmc_adv_kernel_times(g.h, total, capi.dptr(sc[total:total + 4].copy()), 4, capi.dptr(kt), 8,
                                    C.byref(nk))
    if rc:
        raise RuntimeError("gmc_adv_kernel_times failed")
    tgp = nlon * nlat * ntr
    kernels = [{"name": g.lib.gmc_adv_kernel_name(k).decode(), "avg_us": kt[k] * 1e3} for k in range(nk.value)]
    step_bytes = ADV_BYTES_PER_TRACER_GP(ntr) * tgp
    achieved = step_bytes / (ms_step * 1e-3) / 1e9
    result = {
        "metric": "tracer-grid-point-steps/s (FP64 FFSL advection, 1 GPU)",
        "value": tgp / (ms_step * 1e-3),
        "unit": "tracer-grid-point-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "wall_ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded non-divergent wind x cos(pi t / 12 d); cosine bells + one uniform-random tracer)",
        "config": {"workload": f"FFSL {args.config}: {nlon}x{nlat}, {ntr} tracers, FP64, dt={dt:g} s, "
                               f"PPM mono, lon_halo_width {cfg['hw']}",
                   "nlon": nlon, "nlat": nlat, "ntracers": ntr, "dt": dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(args.config, "step"),
                     "scope": f"whole advection step: {ADV_BYTES_PER_TRACER_GP(ntr):.1f} B per tracer-grid-point "
                              "algorithmic; traffic = HBM bytes of one step's launches from the profiler's "
                              "byte counters"},
        "kernels": kernels,
        "mass_rel_drift_bell_tracers": drift,
        "min_tracer": qmin,
        "fixes": g.fixes,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline_adv(st, qs, ub, vb, cfg, args.cpu_budget)
    g.close()
    print(json.dumps(result))


def run_ens(args):
    """Batched ensemble: every member advances `steps` steps; value =
    member-steps/s of the ensemble (one graph per step for all members);
    baseline = the same members run one at a time through gmc_step.  The
    two runs start from the same states and must end bit-identical."""
    from gmcore_amd import init_cond
    from gmcore_amd.host import Ensemble, Setup, ShallowWater
    from gmcore_amd.namelist import Namelist
    cfg = ENS_CONFIGS[args.config]
    steps = args.steps if args.steps_given else cfg["steps"]
    t = time.perf_counter()
    ens_m, plain = [], []
    k = 0
    for nlon, nlat, dt, count in cfg["members"]:
        st = Setup(Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=dt, time_scheme="pc2"))
        for _ in range(count):
            ic = init_cond.seeded_state(nlon, nlat, st.mesh.metrics, cfg["lmax"], 0,
                                        perturb=(cfg["perturb"], 100 + k), omega=st.nl.omega,
                                        radius=st.nl.radius)
            for lst in (ens_m, plain):
                m = ShallowWater(st)
                m.set_state(*ic)
                m.prepare(1)
                lst.append(m)
            k += 1
    log(f"ensemble set-up ({k} members x 2) {time.perf_counter() - t:.1f} s")
    ens = Ensemble(ens_m)
    choice, times = ens.autotune(trial_steps=3)
    log(f"autotune: shape {choice} ({Ensemble.shape_name(choice)}); ms/step "
        + ", ".join(f"{x:.3f}" for x in times))
    # warm-up: one step of every member on both paths (graphs, tables)
    ens.step(1)
    for m in plain:
        m.step(1)
        m.synchronize()
    w0 = time.perf_counter()
    ens.step(steps)
    t_ens = time.perf_counter() - w0
    w0 = time.perf_counter()
    for m in plain:
        m.step(steps)
        m.synchronize()
    t_plain = time.perf_counter() - w0
    same = True
    for a, b in zip(ens_m, plain):
        sa, sb = a.state(), b.state()
        same = same and all(np.array_equal(sa[f], sb[f]) for f in ("u_lon", "v_lat", "gz"))
    n = len(ens_m)
    gp = sum(cnt * nlon * nlat for nlon, nlat, _, cnt in cfg["members"])
    v_ens, v_plain = n * steps / t_ens, n * steps / t_plain
    result = {
        "metric": "member-steps/s (batched FP64 SW ensemble, 1 GPU)",
        "value": v_ens, "unit": "member-steps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": t_ens * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded balanced states, member k perturbed 1 % rms with seed 100 + k)",
        "config": {"workload": f"ensemble {args.config}: " + ", ".join(
            f"{cnt} x {nlon}x{nlat} (dt {dt:g} s)" for nlon, nlat, dt, cnt in cfg["members"]),
            "members": n, "grid_points": gp, "time_scheme": "pc2"},
        "plain_member_steps_per_s": v_plain,
        "speedup_vs_plain": v_ens / v_plain,
        "identical_to_plain": bool(same),
        "shape": Ensemble.shape_name(choice),
        "grid_point_steps_per_s": gp * steps / t_ens,
        "roofline": {"bound": "hbm", "achieved": BYTES_PER_GP_STEP * gp * steps / t_ens / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": BYTES_PER_GP_STEP * gp * steps / t_ens / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "scope": "whole ensemble step, 280 B per member grid point"},
    }
    ens.close()
    return result


def run_adj(args):
    """Gradient of J = 1/2 |gz_N(obs) - y|^2 / sigma_o^2 + 1/2 |x0 - xb|^2 /
    sigma_b^2 over a window of N steps, by the adjoint with checkpointing;
    reports its time against N forward steps of the plain path (value =
    gradient-window time / forward-window time), and the dot-product test of
    the TL and AD over one step at the same size."""
    from gmcore_amd import init_cond
    from gmcore_amd.host import TLAD, Setup, ShallowWater
    from gmcore_amd.namelist import Namelist
    cfg = ADJ_CONFIGS[args.config]
    nlon, nlat, dt, N = cfg["nlon"], cfg["nlat"], cfg["dt"], cfg["window"]
    st = Setup(Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=dt, time_scheme="pc2"))
    m_ = st.mesh.metrics
    truth_ic = init_cond.seeded_state(nlon, nlat, m_, cfg["lmax"], cfg["truth_seed"], omega=st.nl.omega,
                                      radius=st.nl.radius)
    bg = init_cond.seeded_state(nlon, nlat, m_, cfg["lmax"], cfg["truth_seed"], perturb=cfg["bg_perturb"],
                                omega=st.nl.omega, radius=st.nl.radius)
    truth = ShallowWater(st)
    truth.set_state(*truth_ic)
    truth.prepare(1)
    truth.step(N)
    rng = np.random.default_rng(cfg["obs_seed"])
    cells = rng.choice(nlat * nlon, size=cfg["nobs"], replace=False)
    y = truth.state()["gz"].ravel()[cells] + rng.normal(0.0, cfg["sigma_o"], cfg["nobs"])
    truth.close()
    m = ShallowWater(st)
    m.set_state(*bg)
    m.prepare(1)
    t = TLAD(m)
    xb = (bg[0], bg[1], bg[2])
    # warm-up: the same call once (graphs captured, tape and checkpoint
    # buffers allocated), so that the timed call is a repeat of it, as in the
    # iterations of a minimisation, which also keep their gradient arrays
    grad = tuple(np.empty_like(a) for a in xb)
    t.gradient(N, cfg["ckpt"], cells, y, cfg["sigma_o"], xb, cfg["sigma_b"], out=grad)
    m.step(1)
    m.synchronize()
    m.set_state(*bg)
    m.prepare(1)
    w0 = time.perf_counter()
    J, grad = t.gradient(N, cfg["ckpt"], cells, y, cfg["sigma_o"], xb, cfg["sigma_b"], out=grad)
    t_grad = time.perf_counter() - w0
    # N forward steps of the plain path from the same state
    m2 = ShallowWater(st)
    m2.set_state(*bg)
    m2.prepare(1)
    m2.step(1)
    m2.synchronize()
    m2.set_state(*bg)
    m2.prepare(1)
    m2.synchronize()
    w0 = time.perf_counter()
    m2.step(N)
    m2.synchronize()
    t_fwd = time.perf_counter() - w0
    # dot-product test over one step
    rs = np.random.default_rng(7)
    dx = (rs.standard_normal((nlat, nlon)) * 0.5, rs.standard_normal((nlat - 1, nlon)) * 0.5,
          rs.standard_normal((nlat, nlon)) * 100.0)
    dx[0][[0, -1]] = 0.0
    lam = (rs.standard_normal((nlat, nlon)), rs.standard_normal((nlat - 1, nlon)),
           rs.standard_normal((nlat, nlon)) * 1e-2)
    lam[0][[0, -1]] = 0.0
    names = ("u", "v", "gz")
    t.zero()
    lv = m.level
    for k, a in zip(names, dx):
        t.set(lv, k, a)
    t.tl_step(1)
    mdx = [t.get(m.level, k) for k in names]
    m.set_state(*bg, level=m.level)
    m.prepare(m.level)
    t.zero()
    lv = m.level
    for k, a in zip(names, lam):
        t.set(lv, k, a)
    t.ad_step(1)
    mtl = [t.get(lv, k) for k in names]
    dot = lambda a, b: float(np.sum(a[0][1:-1] * b[0][1:-1]) + np.sum(a[1] * b[1]) + np.sum(a[2] * b[2]))
    lhs, rhs = dot(mdx, lam), dot(dx, mtl)
    t.close()
    ratio = t_grad / t_fwd
    result = {
        "metric": "adjoint gradient cost / forward cost (FP64 SW, 1 GPU)",
        "value": ratio, "unit": "x forward window", "n_gpus": 1, "steps": N, "warmup": 1,
        "ms_per_step": t_grad * 1e3 / N, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded truth l<=30; background = truth perturbed 1 % rms; "
                                "5000 observed gz cells with N(0, 50^2) noise)",
        "config": {"workload": f"adjoint {args.config}: {nlon}x{nlat}, pc2, dt={dt:g} s, window {N} steps, "
                               f"checkpoint every {cfg['ckpt']} steps", "nlon": nlon, "nlat": nlat},
        "gradient_ms": t_grad * 1e3, "forward_ms": t_fwd * 1e3, "cost": J,
        "grad_norm": float(np.sqrt(sum(float(np.sum(g * g)) for g in grad))),
        "dot_product_rel_err": abs(lhs - rhs) / max(abs(lhs), abs(rhs)),
    }
    m.close()
    m2.close()
    return result


def run_adv_adj(args):
    """Gradient of J = 1/2 sum (q_N - y)^2 over the tracers of an FFSL window
    (gmc_adv_gradient: window forward with checkpoints, segments recomputed
    with their base states kept, adjoint steps), against the forward of the
    same window (gmc_adv_drive); value = gradient time / forward time.  Both
    are timed on a second window (steps N .. 2N-1) after a first one that
    allocates the checkpoints and tapes; y is the truth run's state at 2N.
    Also the dot-product test of the tangent and adjoint over 4 steps."""
    from gmcore_amd import init_cond
    from gmcore_amd.adv import Advection, AdvSetup
    from gmcore_amd.namelist import Namelist
    acfg = ADV_ADJ_CONFIGS[args.config]
    cfg = ADV_CONFIGS[acfg["base"]]
    nlon, nlat, dt, ntr, N = cfg["nlon"], cfg["nlat"], cfg["dt"], cfg["ntracers"], acfg["window"]
    nl = Namelist(num_lon=nlon, num_lat=nlat, dt_adv=dt, dt_dyn=dt, limiter_type="mono")
    st = AdvSetup(nl, ntr, lon_halo_width=cfg["hw"])
    m = st.mesh.metrics
    ub, vb = init_cond.deform_wind_basis(nlon, nlat, m, cfg["lmax"], cfg["wind_seed"], cfg["vmax"])
    sc = init_cond.wind_scales(2 * N, dt)

    def make(qs):
        g = Advection(st)
        for l, q in enumerate(qs):
            g.set_tracer(1, l, q)
            g.set_tracer(2, l, q)
        g.set_wind_basis(ub, vb)
        return g

    truth = make(init_cond.tracer_set(nlon, nlat, m, cfg["tracer_seed"], ntr))
    truth.drive(0, sc)
    y = np.stack([truth.tracer(truth.old, l) for l in range(ntr)])
    truth.close()
    bg = init_cond.tracer_set(nlon, nlat, m, acfg["bg_seed"], ntr)
    # the gradient array is kept from the first call to the timed one, as a
    # minimisation's iterations keep theirs
    g = make(bg)
    grad = np.empty_like(y)
    g.gradient(0, sc[:N], y, ckpt_every=acfg["ckpt"], out=grad)
    g.synchronize()
    w0 = time.perf_counter()
    J, grad = g.gradient(N, sc[N:], y, ckpt_every=acfg["ckpt"], out=grad)
    t_grad = time.perf_counter() - w0
    g.close()
    # the same gradient with the whole window taped during its forward
    # (checkpoint interval = window: no recompute, N slab copies of memory)
    g = make(bg)
    grad2 = np.empty_like(y)
    g.gradient(0, sc[:N], y, ckpt_every=N, out=grad2)
    g.synchronize()
    w0 = time.perf_counter()
    J2, grad2 = g.gradient(N, sc[N:], y, ckpt_every=N, out=grad2)
    t_grad_full = time.perf_counter() - w0
    g.close()
    same = float(np.abs(grad2 - grad).max() / np.abs(grad).max())
    f = make(bg)
    f.drive(0, sc[:N])
    f.synchronize()
    w0 = time.perf_counter()
    f.drive(N, sc[N:])
    f.synchronize()
    t_fwd = time.perf_counter() - w0
    f.close()
    rs = np.random.default_rng(7)
    d = rs.standard_normal((ntr, nlat, nlon))
    lam = rs.standard_normal((ntr, nlat, nlon))
    a = make(bg)
    tl = a.tl_drive(0, sc[:4], d)
    a.close()
    a = make(bg)
    ad = a.ad_drive(0, sc[:4], lam, ckpt_every=2)
    a.close()
    lhs, rhs = float(np.vdot(lam, tl)), float(np.vdot(ad, d))
    result = {
        "metric": "FFSL adjoint gradient cost / forward cost (FP64, 1 GPU)",
        "value": t_grad / t_fwd, "unit": "x forward window", "n_gpus": 1, "steps": N, "warmup": N,
        "ms_per_step": t_grad * 1e3 / N, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (cosine-bell and random tracers; y = truth run at the window's end, "
                                "background from another seed)",
        "config": {"workload": f"adjoint {args.config}: FFSL {nlon}x{nlat}, {ntr} tracers, dt={dt:g} s, "
                               f"window {N} steps, checkpoint every {acfg['ckpt']} steps", "nlon": nlon,
                   "nlat": nlat, "ntracers": ntr},
        "gradient_ms": t_grad * 1e3, "forward_ms": t_fwd * 1e3, "cost": J,
        "grad_norm": float(np.sqrt(np.sum(grad * grad))),
        "dot_product_rel_err": abs(lhs - rhs) / max(abs(lhs), abs(rhs)),
        "gradient_ms_window_taped": t_grad_full * 1e3, "ratio_window_taped": t_grad_full / t_fwd,
        "window_taped_grad_rel_diff": same,
    }
    return result


def target_figures():
    """The target figures of the two additions, measured in the same run as
    the headline: the SW adjoint gradient's cost over its forward window
    (adj1, target <= 3, dot-product test to 1e-12), the FFSL adjoint's
    (adjffsl05) and the batched ensemble's member-steps/s over the same
    members run one at a time (ens64, target >= 2, identical states).  Each
    is the full line of its own --config, summarised; a failure is reported
    in place of its figures."""
    out = {}
    for name, fn in (("adj1", run_adj), ("adjffsl05", run_adv_adj), ("ens64", run_ens)):
        ns = argparse.Namespace(config=name, steps=0, warmup=1, steps_given=False, no_cpu=True, cpu_budget=0.0)
        try:
            r = fn(ns)
        except Exception as e:   # keep the headline line whatever happens here
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            continue
        if name == "ens64":
            out[name] = {"workload": r["config"]["workload"], "member_steps_per_s": r["value"],
                         "plain_member_steps_per_s": r["plain_member_steps_per_s"],
                         "speedup_vs_plain": r["speedup_vs_plain"], "identical_to_plain": r["identical_to_plain"],
                         "target": "speedup >= 2, identical results"}
        else:
            out[name] = {"workload": r["config"]["workload"], "gradient_over_forward": r["value"],
                         "gradient_ms": r["gradient_ms"], "forward_ms": r["forward_ms"],
                         "dot_product_rel_err": r["dot_product_rel_err"],
                         "target": "<= 3, dot-product test to 1e-12" if name == "adj1"
                                   else "dot-product test to 1e-12"}
            if "ratio_window_taped" in r:
                out[name]["gradient_over_forward_window_taped"] = r["ratio_window_taped"]
    return out


def run_baro(args):
    """The hydrostatic column operators on a 1-degree, 32-level block: each
    operator launched `reps` times back to back between two HIP events on
    the block's stream; GB/s of its algorithmic bytes.  value = column-level
    points per second of the sequence calc_ph, calc_t, calc_gz_lev,
    pgf_lin97 (the hydrostatic pressure gradient of a backward pass)."""
    import ctypes as C
    import oracle
    from gmcore_amd.baro import CPD, P0, RD, Baroclinic
    from gmcore_amd.host import Filter, Mesh
    from gmcore_amd.namelist import Namelist
    cfg = BARO_CONFIGS[args.config]
    nlon, nlat, nlev, reps = cfg["nlon"], cfg["nlat"], cfg["nlev"], cfg["reps"]
    mesh = Mesh(nlon, nlat, 6371220.0, 7.292e-5)
    hyai, hybi = oracle.equal_eta(nlev)
    g = Baroclinic(mesh, hyai, hybi, P0, cfg["ptop"], RD, CPD, P0)
    rng = np.random.default_rng(0)
    col = np.linspace(1.0, 0.0, nlev)[:, None, None]
    g.set("phs", 1.0e5 + 2000.0 * rng.standard_normal((1, nlat, nlon)))
    g.set("pt", 300.0 + 60.0 * col + rng.standard_normal((nlev, nlat, nlon)))
    gzs = 2000.0 * rng.random((1, nlat, nlon))
    g.set("gzs", gzs)
    bottom = np.zeros((nlev + 1, nlat, nlon))
    bottom[-1] = gzs[0]
    g.set("gz_lev", bottom)
    for name in ("dmfdlon", "dmfdlat"):
        g.set(name, 1e-3 * rng.standard_normal((nlev, nlat, nlon)))
    g.set("u_f", 20 * rng.standard_normal((nlev, nlat, nlon)))
    g.set("v_f", 20 * rng.standard_normal((nlev, nlat - 1, nlon)))
    g.set("u_lon", 20 * rng.standard_normal((nlev, nlat, nlon)))
    g.set("v_lat", 20 * rng.standard_normal((nlev, nlat - 1, nlon)))
    g.set("m_lon", 1e3 + 4e3 * rng.random((nlev, nlat, nlon)))
    g.set("m_lat", 1e3 + 4e3 * rng.random((nlev, nlat - 1, nlon)))
    c = rng.uniform(-0.9, 0.9, (nlev + 1, nlat, nlon))
    g.set("cflz", c)
    # the mass batch of calc_grad_ptf's horizontal fluxes: |cflx| up to 2.9
    # (whole cells crossed) within lon_halo_width 3, |cfly| < 0.9
    shp, shv = (nlev, nlat, nlon), (nlev, nlat - 1, nlon)
    for name, a in (("adv_u", 20 * rng.standard_normal(shp)), ("adv_v", 20 * rng.standard_normal(shv)),
                    ("adv_mfx", 1e4 * rng.standard_normal(shp)), ("adv_mfy", 1e4 * rng.standard_normal(shv)),
                    ("adv_cflx", rng.uniform(-2.9, 2.9, shp)), ("adv_cfly", rng.uniform(-0.9, 0.9, shv)),
                    ("adv_divx", 1e-5 * rng.standard_normal(shp)), ("adv_divy", 1e-5 * rng.standard_normal(shp))):
        g.set(name, a)
    table = Filter(mesh, 300.0, Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=300.0), 2)
    ops = [("calc_ph", g.calc_ph), ("calc_t", g.calc_t), ("calc_gz_lev", g.calc_gz_lev),
           ("calc_dphsdt", g.calc_dphsdt), ("calc_we_lev", g.calc_we_lev),
           ("calc_wedudlev_wedvdlev", g.calc_wedudlev_wedvdlev), ("pgf_lin97", g.pgf_lin97),
           ("calc_ptf_lev", lambda: g.calc_ptf_lev(check=False)),
           ("calc_ptf_hflx", lambda: g.calc_ptf_hflx(300.0, lon_halo_width=3, check=False)),
           ("small_filter2", g.small_filter2)]
    g.set_small_filter2(table)
    hip = C.CDLL("libamdhip64.so")
    stream = C.c_void_p()
    assert g.lib.gmc_baro_stream(g.h, C.byref(stream)) == 0
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    for _, f in ops:
        f()
    g.synchronize()
    cols = nlon * nlat
    kernels = []
    for name, f in ops:
        best = 1e30
        for _ in range(3):
            hip.hipEventRecord(e0, stream)
            for _ in range(reps):
                f()
            hip.hipEventRecord(e1, stream)
            hip.hipEventSynchronize(e1)
            ms = C.c_float()
            hip.hipEventElapsedTime(C.byref(ms), e0, e1)
            best = min(best, ms.value / reps)
        full, half, one = BARO_FIELDS_PER_OP[name]
        nbytes = 8.0 * cols * (full * nlev + half * (nlev + 1) + one)
        kernels.append({"name": name, "avg_us": best * 1e3, "bytes": nbytes,
                        "gbs": nbytes / (best * 1e-3) / 1e9})
    g.synchronize()   # raises if a timed launch met a Courant number out of range
    seq = ("calc_ph", "calc_t", "calc_gz_lev", "pgf_lin97")
    t_seq = sum(k["avg_us"] for k in kernels if k["name"] in seq)
    dom = max((k for k in kernels if k["name"] in seq), key=lambda k: k["avg_us"])
    result = {
        "metric": "column-level-points/s (FP64 hydrostatic column operators, 1 GPU)",
        "value": cols * nlev / (t_seq * 1e-6),
        "unit": "column-level-points/s",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": t_seq * 1e-3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (equal-eta coordinate, stratified pt, random fluxes and winds)",
        "config": {"workload": f"baroclinic column operators {args.config}: {nlon}x{nlat}x{nlev} levels, FP64",
                   "nlon": nlon, "nlat": nlat, "nlev": nlev},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": profiled_traffic(args.config, dom["name"]),
                     "kernel": dom["name"],
                     "scope": "slowest operator of the timed sequence: algorithmic bytes over its mean time in "
                              "back-to-back launches"},
        "kernels": kernels,
    }
    print(json.dumps(result))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="sw025",
                    choices=sorted(CONFIGS) + sorted(ADV_CONFIGS) + sorted(ENS_CONFIGS) + sorted(ADJ_CONFIGS)
                    + sorted(ADV_ADJ_CONFIGS) + sorted(BARO_CONFIGS))
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg (and the additions' target figures: the GPU leg of the "
                         "headline workload only, as profiling runs want it)")
    ap.add_argument("--no-targets", action="store_true",
                    help="skip the additions' target figures of the default line")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB five times (for scaling the HBM "
                         "byte counters of a profiler against a kernel of known traffic)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 8 * n, "copy_calibration_bytes_written": 8 * n,
                          "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU latitude-band decomposition is a design, not code)", file=sys.stderr)
        sys.exit(2)
    args.steps_given = "--steps" in sys.argv
    if args.config in ADV_CONFIGS:
        run_adv(args)
        return
    if args.config in ENS_CONFIGS:
        print(json.dumps(run_ens(args)))
        return
    if args.config in ADJ_CONFIGS:
        print(json.dumps(run_adj(args)))
        return
    if args.config in ADV_ADJ_CONFIGS:
        print(json.dumps(run_adv_adj(args)))
        return
    if args.config in BARO_CONFIGS:
        run_baro(args)
        return
    cfg = CONFIGS[args.config]
    from gmcore_amd.host import ShallowWater
    from gmcore_amd import capi
    import ctypes as C

    t = time.perf_counter()
    st = make_setup(cfg)
    ic = make_ic(st, cfg)
    log(f"setup + initial state {time.perf_counter() - t:.2f} s")
    m = ShallowWater(st)
    m.set_state(*ic)
    m.prepare(1)
    d0 = m.diagnose(1)
    m.step(args.warmup)
    m.synchronize()

    # timed region: HIP events on the library's stream + host wall clock
    hip = C.CDLL("libamdhip64.so")
    stream = C.c_void_p()
    assert m.lib.gmc_stream(m.ctx, C.byref(stream)) == 0
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(ev0)) == 0 and hip.hipEventCreate(C.byref(ev1)) == 0
    m.synchronize()
    w0 = time.perf_counter()
    hip.hipEventRecord(ev0, stream)
    m.step(args.steps)
    hip.hipEventRecord(ev1, stream)
    hip.hipEventSynchronize(ev1)
    m.synchronize()
    wall = time.perf_counter() - w0
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    ms_step = ms.value / args.steps
    d1 = m.diagnose()

    # per-kernel breakdown (launches one by one with events)
    kt = np.zeros(8)
    nk = C.c_int32()
    rc = m.lib.gmc_kernel_times(m.ctx, cfg["dt"], 3, capi.dptr(kt), 8, C.byref(nk))
    if rc:
        raise RuntimeError("gmc_kernel_times failed")
    gp = cfg["nlon"] * cfg["nlat"]
    kernels = []
    for k in range(nk.value):
        name = m.lib.gmc_kernel_name(k).decode()
        b = KERNEL_BYTES_PER_GP.get(name, 0.0) * gp
        kernels.append({"name": name, "avg_us": kt[k] * 1e3,
                        "bytes_per_launch": b, "gbs": b / (kt[k] * 1e-3) / 1e9 if kt[k] > 0 else 0.0})
    dom = max(kernels, key=lambda x: x["avg_us"])
    step_gbs = BYTES_PER_GP_STEP * gp / (ms_step * 1e-3) / 1e9
    traffic = profiled_traffic(args.config, dom["name"])
    result = {
        "metric": "grid-point-steps/s (FP64 SW pc2 step, 1 GPU)",
        "value": gp / (ms_step * 1e-3),
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "wall_ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded spherical-harmonic balanced initial state)",
        "config": {"workload": f"shallow-water {args.config}: {cfg['nlon']}x{cfg['nlat']} C-grid, FP64, "
                               f"pc2, dt={cfg['dt']:g} s, polar filter on, seeded l<={cfg['lmax']}",
                   "nlon": cfg["nlon"], "nlat": cfg["nlat"], "dt": cfg["dt"], "time_scheme": "pc2"},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": dom["name"],
                     "scope": f"dominant kernel {dom['name']}: algorithmic bytes per launch "
                              f"{KERNEL_BYTES_PER_GP[dom['name']]:.2f} B/grid point over its mean "
                              "launch time (start and end HIP events of each launch on the library stream)"},
        "step_roofline": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS,
                          "scope": "whole pc2 step (one graph launch): 280 B/grid point algorithmic"},
        "dominant_kernel": dom,
        "kernels": kernels,
        "mass_rel_drift": float((d1[0] - d0[0]) / d0[0]),
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(st, ic, cfg, args.cpu_budget)
        result["cpu_baseline_all_cores"] = cpu_baseline(st, ic, cfg, args.cpu_budget, omp=True)
    m.close()
    if not args.no_cpu and not args.no_targets:
        log("target figures of the additions (adj1, adjffsl05, ens64)")
        result["targets"] = target_figures()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
