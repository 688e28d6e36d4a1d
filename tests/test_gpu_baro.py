"""The hydrostatic column operators of libgmcore_hip.so (gmc_baro_*) against
the oracle (oracle/baro.c) on the same inputs: bit for bit where the
operator is additions, products and divisions; within 1e-15 relative where
it calls pow, 1e-13 where it sums logs down a column, and 1e-12 relative
(the per-step bar of FP64 fields) for the chained column state."""
import numpy as np
import pytest

import oracle
from gmcore_amd.baro import CPD, P0, RD, Baroclinic
from gmcore_amd.host import Mesh

pytestmark = pytest.mark.gpu

NLON, NLAT, NLEV = 180, 91, 26
PTOP = 219.4          # namelist ptop (namelist_mod.F90:53), the equal_eta template's local_ptop


@pytest.fixture(scope="module")
def env():
    mesh = Mesh(NLON, NLAT, 6371220.0, 7.292e-5)
    hyai, hybi = oracle.equal_eta(NLEV)
    o = oracle.BaroOracle(mesh.metrics, NLON, NLAT, hyai, hybi, P0, PTOP, RD, CPD, P0)
    return mesh, hyai, hybi, o


def gpu(env):
    mesh, hyai, hybi, _ = env
    return Baroclinic(mesh, hyai, hybi, P0, PTOP, RD, CPD, P0)


def rng(seed):
    return np.random.default_rng(seed)


def surface_pressure(seed):
    return 1.0e5 + 3000.0 * rng(seed).standard_normal((1, NLAT, NLON))


def column_state(o, seed):
    """A realistic column state: pressures of a perturbed surface pressure,
    a stratified potential temperature, hilly surface geopotential."""
    r = rng(seed)
    phs = surface_pressure(seed)
    ph_lev, exn, ph = o.calc_ph(phs)
    theta = 300.0 + 60.0 * np.linspace(1.0, 0.0, NLEV)[:, None, None] + r.standard_normal((NLEV, NLAT, NLON))
    gzs = 2000.0 * r.random((1, NLAT, NLON))
    return phs, ph_lev, exn, ph, theta, gzs


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_calc_ph(env):
    o = env[3]
    g = gpu(env)
    phs = surface_pressure(1)
    g.set("phs", phs)
    g.calc_ph()
    ph_lev, exn, ph = o.calc_ph(phs)
    np.testing.assert_array_equal(g.get("ph_lev"), ph_lev)
    np.testing.assert_array_equal(g.get("ph"), ph)
    np.testing.assert_allclose(g.get("ph_exn_lev"), exn, rtol=1e-15, atol=0)


def test_calc_t(env):
    o = env[3]
    g = gpu(env)
    _, _, _, ph, theta, _ = column_state(o, 2)
    g.set("pt", theta)
    g.set("ph", ph)
    g.calc_t()
    np.testing.assert_allclose(g.get("t"), o.calc_t(theta, ph), rtol=1e-15, atol=0)


def test_calc_gz_lev(env):
    o = env[3]
    g = gpu(env)
    _, ph_lev, _, ph, theta, gzs = column_state(o, 3)
    t = o.calc_t(theta, ph)
    bottom = np.zeros((NLEV + 1, NLAT, NLON))
    bottom[-1] = gzs[0]
    for name, a in (("gzs", gzs), ("t", t), ("ph_lev", ph_lev), ("gz_lev", bottom)):
        g.set(name, a)
    g.calc_gz_lev()
    gz_lev, gz = o.calc_gz_lev(gzs, t, ph_lev, bottom)
    got = g.get("gz_lev")
    np.testing.assert_array_equal(got[-1], gzs[0])
    assert rel(got, gz_lev) < 1e-13
    assert rel(g.get("gz"), gz) < 1e-13


def test_dphsdt_and_we_lev_bitwise(env):
    o = env[3]
    g = gpu(env)
    r = rng(4)
    a, b = 1e-3 * r.standard_normal((NLEV, NLAT, NLON)), 1e-3 * r.standard_normal((NLEV, NLAT, NLON))
    g.set("dmfdlon", a)
    g.set("dmfdlat", b)
    g.calc_dphsdt()
    dphs = o.calc_dphsdt(a, b)
    np.testing.assert_array_equal(g.get("dphs"), dphs)
    g.calc_we_lev()
    we, wlon, wlat = o.calc_we_lev(a, b, dphs)
    np.testing.assert_array_equal(g.get("we_lev"), we)
    np.testing.assert_array_equal(g.get("we_lev_lon"), wlon)
    np.testing.assert_array_equal(g.get("we_lev_lat"), wlat)


def test_wedudlev_wedvdlev_bitwise(env):
    o = env[3]
    g = gpu(env)
    r = rng(5)
    u, v = 30 * r.standard_normal((NLEV, NLAT, NLON)), 30 * r.standard_normal((NLEV, NLAT - 1, NLON))
    m_lon, m_lat = 1e3 + 4e3 * r.random((NLEV, NLAT, NLON)), 1e3 + 4e3 * r.random((NLEV, NLAT - 1, NLON))
    w_lon, w_lat = r.standard_normal((NLEV + 1, NLAT, NLON)), r.standard_normal((NLEV + 1, NLAT - 1, NLON))
    for name, x in (("u_f", u), ("v_f", v), ("m_lon", m_lon), ("m_lat", m_lat), ("we_lev_lon", w_lon),
                    ("we_lev_lat", w_lat)):
        g.set(name, x)
    g.calc_wedudlev_wedvdlev()
    du, dv = o.calc_wedudlev_wedvdlev(u, v, m_lon, m_lat, w_lon, w_lat)
    np.testing.assert_array_equal(g.get("wedudlev"), du)
    np.testing.assert_array_equal(g.get("wedvdlev"), dv)


def test_pgf_lin97_bitwise(env):
    o = env[3]
    g = gpu(env)
    _, ph_lev, exn, ph, theta, gzs = column_state(o, 6)
    gz_lev, _ = o.calc_gz_lev(gzs, o.calc_t(theta, ph), ph_lev)
    gz_lev[-1] = gzs[0]
    g.set("ph_exn_lev", exn)
    g.set("gz_lev", gz_lev)
    g.pgf_lin97()
    pl, pt = o.pgf_lin97(exn, gz_lev)
    np.testing.assert_array_equal(g.get("pgf_lon"), pl)
    np.testing.assert_array_equal(g.get("pgf_lat"), pt)


def test_column_chain(env):
    """backward_pass's hydrostatic prelude (operators_mod.F90:124-128) and the
    pressure gradient: calc_ph -> calc_t -> calc_gz_lev -> pgf_lin97 on the
    device, each on the previous one's device output, against the same chain
    on the CPU."""
    o = env[3]
    g = gpu(env)
    phs, _, _, _, theta, gzs = column_state(o, 7)
    bottom = np.zeros((NLEV + 1, NLAT, NLON))
    bottom[-1] = gzs[0]
    for name, a in (("phs", phs), ("pt", theta), ("gzs", gzs), ("gz_lev", bottom)):
        g.set(name, a)
    g.calc_ph()
    g.calc_t()
    g.calc_gz_lev()
    g.pgf_lin97()
    ph_lev, exn, ph = o.calc_ph(phs)
    t = o.calc_t(theta, ph)
    gz_lev, gz = o.calc_gz_lev(gzs, t, ph_lev, bottom)
    pl, pt = o.pgf_lin97(exn, gz_lev)
    for name, want in (("t", t), ("gz_lev", gz_lev), ("gz", gz), ("pgf_lon", pl), ("pgf_lat", pt)):
        assert rel(g.get(name), want) < 1e-12, name


def test_rejects_bad_arguments(env):
    mesh, hyai, hybi, _ = env
    with pytest.raises(Exception):
        Baroclinic(mesh, hyai[:2], hybi[:2])     # one level
    g = gpu(env)
    with pytest.raises(Exception):
        g.set("pt", np.zeros((NLEV - 1, NLAT, NLON)))


def courant_numbers(seed):
    r = rng(seed)
    c = r.uniform(-0.95, 0.95, (NLEV + 1, NLAT, NLON))
    c[3:NLEV - 2] = r.uniform(-2.9, 2.9, (NLEV - 5, NLAT, NLON))
    return c


@pytest.mark.parametrize("limiter", ["mono", "pd", "none"])
def test_calc_ptf_lev_bitwise(env, limiter):
    o = env[3]
    g = gpu(env)
    _, _, _, _, theta, _ = column_state(o, 8)
    w = rng(9).standard_normal((NLEV + 1, NLAT, NLON))
    c = courant_numbers(10)
    for name, a in (("pt", theta), ("we_lev", w), ("cflz", c)):
        g.set(name, a)
    g.calc_ptf_lev(limiter)
    ptf, dptf, bad = o.calc_ptf_lev(theta, w, c, {"mono": 0, "pd": 1, "none": 2}[limiter])
    assert bad == 0
    np.testing.assert_array_equal(g.get("ptf_lev"), ptf)
    np.testing.assert_array_equal(g.get("dptfdlev"), dptf)


def test_calc_ptf_lev_reports_courant_beyond_column(env):
    """An upwind cell above the model top is refused (GMC_ENUMERIC) with every
    access kept inside the arrays."""
    g = gpu(env)
    c = np.zeros((NLEV + 1, NLAT, NLON))
    c[1, 5, 7] = 1.5
    g.set("pt", np.full((NLEV, NLAT, NLON), 300.0))
    g.set("we_lev", np.ones_like(c))
    g.set("cflz", c)
    with pytest.raises(Exception, match="outside the column"):
        g.calc_ptf_lev()
    g.synchronize()                   # reported once


def test_calc_ptf_lev_error_reported_by_synchronize(env):
    """Without the check the call only queues the work: the next synchronize
    reports the Courant number below the model bottom, the one after is
    clean, and a clean launch in between does not hide it."""
    g = gpu(env)
    c = np.zeros((NLEV + 1, NLAT, NLON))
    c[NLEV - 1, 3, 4] = -1.5
    g.set("pt", np.full((NLEV, NLAT, NLON), 300.0))
    g.set("we_lev", np.ones_like(c))
    g.set("cflz", c)
    g.calc_ptf_lev(check=False)
    g.set("cflz", np.zeros_like(c))
    g.calc_ptf_lev(check=False)
    with pytest.raises(Exception, match="outside the column"):
        g.synchronize()
    g.synchronize()
    g.calc_ptf_lev()


def test_small_filter2_bitwise(env):
    from gmcore_amd.host import Filter
    from gmcore_amd.namelist import Namelist
    mesh, _, _, o = env
    nl = Namelist(num_lon=NLON, num_lat=NLAT, dt_dyn=600.0)
    table = Filter(mesh, 600.0, nl, 2)
    g = gpu(env)
    r = rng(11)
    u, v = 20 * r.standard_normal((NLEV, NLAT, NLON)), 20 * r.standard_normal((NLEV, NLAT - 1, NLON))
    g.set("u_lon", u)
    g.set("v_lat", v)
    g.small_filter2(table)
    fu, fv = o.small_filter2(u, v, (table.ngrid_lon, table.wgt_lon, table.ngrid_lat, table.wgt_lat))
    np.testing.assert_array_equal(g.get("u_lon"), fu)
    np.testing.assert_array_equal(g.get("v_lat"), fv)
    assert not np.array_equal(fu, u)


def test_small_filter2_needs_a_table(env):
    g = gpu(env)
    with pytest.raises(Exception, match="no small_filter2 table"):
        g.small_filter2()


def test_small_filter2_table_set_once(env):
    """The table set once serves every later call, as filter_init's does."""
    from gmcore_amd.host import Filter
    from gmcore_amd.namelist import Namelist
    mesh, _, _, o = env
    table = Filter(mesh, 600.0, Namelist(num_lon=NLON, num_lat=NLAT, dt_dyn=600.0), 2)
    g = gpu(env)
    g.set_small_filter2(table)
    r = rng(12)
    u, v = 20 * r.standard_normal((NLEV, NLAT, NLON)), 20 * r.standard_normal((NLEV, NLAT - 1, NLON))
    g.set("u_lon", u)
    g.set("v_lat", v)
    g.small_filter2()
    g.small_filter2()
    t = (table.ngrid_lon, table.wgt_lon, table.ngrid_lat, table.wgt_lat)
    fu, fv = o.small_filter2(*o.small_filter2(u, v, t), t)
    np.testing.assert_array_equal(g.get("u_lon"), fu)
    np.testing.assert_array_equal(g.get("v_lat"), fv)


@pytest.mark.parametrize("nlev", [2, 5, 137, 300])
def test_level_counts(nlev):
    """Level counts off the chunk size (2, 5), a 137-level column (the
    geopotential sums' LDS block narrows to 16 columns) and 300 levels (8
    columns): the hydrostatic chain and the vertical flux bit for bit or to
    the chain's bar on a small mesh."""
    nlon, nlat = 36, 19
    mesh = Mesh(nlon, nlat, 6371220.0, 7.292e-5)
    hyai, hybi = oracle.equal_eta(nlev)
    o = oracle.BaroOracle(mesh.metrics, nlon, nlat, hyai, hybi, P0, PTOP, RD, CPD, P0)
    g = Baroclinic(mesh, hyai, hybi, P0, PTOP, RD, CPD, P0)
    r = rng(13)
    phs = 1.0e5 + 3000.0 * r.standard_normal((1, nlat, nlon))
    theta = 300.0 + 60.0 * np.linspace(1.0, 0.0, nlev)[:, None, None] + r.standard_normal((nlev, nlat, nlon))
    gzs = 2000.0 * r.random((1, nlat, nlon))
    bottom = np.zeros((nlev + 1, nlat, nlon))
    bottom[-1] = gzs[0]
    for name, a in (("phs", phs), ("pt", theta), ("gzs", gzs), ("gz_lev", bottom)):
        g.set(name, a)
    g.calc_ph()
    g.calc_t()
    g.calc_gz_lev()
    g.pgf_lin97()
    ph_lev, exn, ph = o.calc_ph(phs)
    np.testing.assert_array_equal(g.get("ph_lev"), ph_lev)
    np.testing.assert_array_equal(g.get("ph"), ph)
    t = o.calc_t(theta, ph)
    gz_lev, gz = o.calc_gz_lev(gzs, t, ph_lev, bottom)
    pl, pt = o.pgf_lin97(exn, gz_lev)
    for name, want in (("t", t), ("gz_lev", gz_lev), ("gz", gz), ("pgf_lon", pl), ("pgf_lat", pt)):
        assert rel(g.get(name), want) < 1e-12, name
    # the geopotential sums alone, on the oracle's own temperature: the
    # ascending order of every half level's sum kept
    g.set("t", t)
    g.set("ph_lev", ph_lev)
    g.calc_gz_lev()
    assert rel(g.get("gz_lev"), gz_lev) < 1e-13
    w = r.standard_normal((nlev + 1, nlat, nlon))
    c = r.uniform(-0.95, 0.95, (nlev + 1, nlat, nlon))
    for name, a in (("pt", theta), ("we_lev", w), ("cflz", c)):
        g.set(name, a)
    g.calc_ptf_lev()
    ptf, dptf, bad = o.calc_ptf_lev(theta, w, c, 0)
    assert bad == 0
    np.testing.assert_array_equal(g.get("ptf_lev"), ptf)
    np.testing.assert_array_equal(g.get("dptfdlev"), dptf)


HW = 3


def mass_batch(seed):
    """Random mass-batch fields: |cflx| up to 2.9 (whole cells crossed),
    |cfly| < 0.9."""
    r = rng(seed)
    shp, shv = (NLEV, NLAT, NLON), (NLEV, NLAT - 1, NLON)
    return {"adv_u": 20 * r.standard_normal(shp), "adv_v": 20 * r.standard_normal(shv),
            "adv_mfx": 1e4 * r.standard_normal(shp), "adv_mfy": 1e4 * r.standard_normal(shv),
            "adv_cflx": r.uniform(-2.9, 2.9, shp), "adv_cfly": r.uniform(-0.9, 0.9, shv),
            "adv_divx": 1e-5 * r.standard_normal(shp), "adv_divy": 1e-5 * r.standard_normal(shp)}


def oracle_hflx(o, pt, f, dt, limiter, flux_type):
    return o.calc_ptf_hflx(pt, f["adv_u"], f["adv_v"], f["adv_mfx"], f["adv_mfy"], f["adv_cflx"], f["adv_cfly"],
                           f["adv_divx"], f["adv_divy"], dt, limiter, flux_type, HW)


@pytest.mark.parametrize("limiter,flux_type", [("mono", "ppm"), ("pd", "ppm"), ("none", "ppm"),
                                               ("mono", "van_leer")])
def test_calc_ptf_hflx_bitwise(env, limiter, flux_type):
    """calc_grad_ptf's horizontal fluxes and their divergence, every level,
    bit for bit against the oracle (pole rows through the zonal sums)."""
    o = env[3]
    g = gpu(env)
    _, _, _, _, theta, _ = column_state(o, 14)
    f = mass_batch(15)
    g.set("pt", theta)
    for k, a in f.items():
        g.set(k, a)
    g.calc_ptf_hflx(600.0, limiter, flux_type, HW)
    want = oracle_hflx(o, theta, f, 600.0, {"mono": 0, "pd": 1, "none": 2}[limiter],
                       {"ppm": 0, "van_leer": 1}[flux_type])
    for name, w in zip(("ptf_lon", "ptf_lat", "dptfdlon", "dptfdlat"), want):
        np.testing.assert_array_equal(g.get(name), w, err_msg=name)


def test_calc_grad_ptf_is_both_parts(env):
    """calc_grad_ptf = the horizontal fluxes, then the vertical ones, each
    against the oracle."""
    o = env[3]
    g = gpu(env)
    _, _, _, _, theta, _ = column_state(o, 16)
    f = mass_batch(17)
    w = rng(18).standard_normal((NLEV + 1, NLAT, NLON))
    c = courant_numbers(19)
    g.set("pt", theta)
    for k, a in list(f.items()) + [("we_lev", w), ("cflz", c)]:
        g.set(k, a)
    g.calc_grad_ptf(300.0, "mono", "ppm", HW)
    want = oracle_hflx(o, theta, f, 300.0, 0, 0)
    for name, x in zip(("ptf_lon", "ptf_lat", "dptfdlon", "dptfdlat"), want):
        np.testing.assert_array_equal(g.get(name), x, err_msg=name)
    ptf, dptf, bad = o.calc_ptf_lev(theta, w, c, 0)
    assert bad == 0
    np.testing.assert_array_equal(g.get("ptf_lev"), ptf)
    np.testing.assert_array_equal(g.get("dptfdlev"), dptf)


def test_calc_ptf_hflx_reports_cflx_beyond_halo(env):
    """A cflx beyond lon_halo_width is reported at the next synchronize (the
    reference stops there); the accesses stay in the row meanwhile."""
    g = gpu(env)
    f = mass_batch(20)
    f["adv_cflx"][4, 10, 11] = HW + 0.25
    g.set("pt", np.full((NLEV, NLAT, NLON), 300.0))
    for k, a in f.items():
        g.set(k, a)
    g.calc_ptf_hflx(600.0, lon_halo_width=HW, check=False)
    with pytest.raises(Exception, match="cflx exceeds lon_halo_width"):
        g.synchronize()
    g.synchronize()
    with pytest.raises(Exception, match="lon_halo_width"):
        g.calc_ptf_hflx(600.0, lon_halo_width=NLON // 2)   # refused before any launch


def test_every_operator_on_a_grid_whose_rows_straddle_wavefronts():
    """72 x 37 (row pitch 96 doubles: a wavefront spans two latitude rows, so
    the U/V branches of a cell's thread and the row tests diverge inside a
    wavefront) with 7 levels (a chunk of 3): every operator against the
    oracle, bit for bit or to its stated bar."""
    from gmcore_amd.host import Filter
    from gmcore_amd.namelist import Namelist
    nlon, nlat, nlev, hw = 72, 37, 7, 3
    mesh = Mesh(nlon, nlat, 6371220.0, 7.292e-5)
    hyai, hybi = oracle.equal_eta(nlev)
    o = oracle.BaroOracle(mesh.metrics, nlon, nlat, hyai, hybi, P0, PTOP, RD, CPD, P0)
    g = Baroclinic(mesh, hyai, hybi, P0, PTOP, RD, CPD, P0)
    r = rng(30)
    shp, shv, shh = (nlev, nlat, nlon), (nlev, nlat - 1, nlon), (nlev + 1, nlat, nlon)
    phs = 1.0e5 + 3000.0 * r.standard_normal((1, nlat, nlon))
    theta = 300.0 + 60.0 * np.linspace(1.0, 0.0, nlev)[:, None, None] + r.standard_normal(shp)
    gzs = 2000.0 * r.random((1, nlat, nlon))
    bottom = np.zeros(shh)
    bottom[-1] = gzs[0]
    for name, a in (("phs", phs), ("pt", theta), ("gzs", gzs), ("gz_lev", bottom)):
        g.set(name, a)
    g.calc_ph()
    g.calc_t()
    g.calc_gz_lev()
    g.pgf_lin97()
    ph_lev, exn, ph = o.calc_ph(phs)
    t = o.calc_t(theta, ph)
    gz_lev, gz = o.calc_gz_lev(gzs, t, ph_lev, bottom)
    np.testing.assert_array_equal(g.get("ph_lev"), ph_lev)
    for name, want in (("t", t), ("gz_lev", gz_lev), ("gz", gz)):
        assert rel(g.get(name), want) < 1e-12, name
    g.set("ph_exn_lev", exn)
    g.set("gz_lev", gz_lev)
    g.pgf_lin97()
    pl, pt_ = o.pgf_lin97(exn, gz_lev)
    np.testing.assert_array_equal(g.get("pgf_lon"), pl)
    np.testing.assert_array_equal(g.get("pgf_lat"), pt_)
    a, b = 1e-3 * r.standard_normal(shp), 1e-3 * r.standard_normal(shp)
    g.set("dmfdlon", a)
    g.set("dmfdlat", b)
    g.calc_dphsdt()
    g.calc_we_lev()
    dphs = o.calc_dphsdt(a, b)
    we, wlon, wlat = o.calc_we_lev(a, b, dphs)
    for name, want in (("dphs", dphs), ("we_lev", we), ("we_lev_lon", wlon), ("we_lev_lat", wlat)):
        np.testing.assert_array_equal(g.get(name), want, err_msg=name)
    u, v = 30 * r.standard_normal(shp), 30 * r.standard_normal(shv)
    ml, mt = 1e3 + 4e3 * r.random(shp), 1e3 + 4e3 * r.random(shv)
    for name, x in (("u_f", u), ("v_f", v), ("m_lon", ml), ("m_lat", mt)):
        g.set(name, x)
    g.calc_wedudlev_wedvdlev()
    du, dv = o.calc_wedudlev_wedvdlev(u, v, ml, mt, wlon, wlat)
    np.testing.assert_array_equal(g.get("wedudlev"), du)
    np.testing.assert_array_equal(g.get("wedvdlev"), dv)
    f = {"adv_u": 20 * r.standard_normal(shp), "adv_v": 20 * r.standard_normal(shv),
         "adv_mfx": 1e4 * r.standard_normal(shp), "adv_mfy": 1e4 * r.standard_normal(shv),
         "adv_cflx": r.uniform(-2.9, 2.9, shp), "adv_cfly": r.uniform(-0.9, 0.9, shv),
         "adv_divx": 1e-5 * r.standard_normal(shp), "adv_divy": 1e-5 * r.standard_normal(shp)}
    w = r.standard_normal(shh)
    c = r.uniform(-0.95, 0.95, shh)
    for k, x in list(f.items()) + [("we_lev", w), ("cflz", c)]:
        g.set(k, x)
    g.calc_grad_ptf(600.0, "pd", "ppm", hw)
    want = o.calc_ptf_hflx(theta, f["adv_u"], f["adv_v"], f["adv_mfx"], f["adv_mfy"], f["adv_cflx"],
                           f["adv_cfly"], f["adv_divx"], f["adv_divy"], 600.0, 1, 0, hw)
    for name, x in zip(("ptf_lon", "ptf_lat", "dptfdlon", "dptfdlat"), want):
        np.testing.assert_array_equal(g.get(name), x, err_msg=name)
    ptf, dptf, bad = o.calc_ptf_lev(theta, w, c, 1)
    assert bad == 0
    np.testing.assert_array_equal(g.get("ptf_lev"), ptf)
    np.testing.assert_array_equal(g.get("dptfdlev"), dptf)
    table = Filter(mesh, 3600.0, Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=3600.0), 2)
    u2, v2 = 20 * r.standard_normal(shp), 20 * r.standard_normal(shv)
    g.set("u_lon", u2)
    g.set("v_lat", v2)
    g.small_filter2(table)
    fu, fv = o.small_filter2(u2, v2, (table.ngrid_lon, table.wgt_lon, table.ngrid_lat, table.wgt_lat))
    np.testing.assert_array_equal(g.get("u_lon"), fu)
    np.testing.assert_array_equal(g.get("v_lat"), fv)
