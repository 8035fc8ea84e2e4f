"""Closed-form checks of the Chem_Mie query and the per-tracer optics sum,
shared by the oracle test (CPU) and the library test (GPU).

The checks follow from the mathematics of the reference's algorithm, not from
a restatement of its formulas:

* node values: a table whose RH grid is RH_i = 20 i / 1000 (exactly the RH
  map's abscissae) queried at RH_i + 0.0005 snaps to map slot 20 i, whose
  interpolation weight is 0, so the query returns the node's own bext, bsca,
  g and moments: tau = bext_i q and ssa = bsca_i / bext_i as single float32
  operations (CB/Chem_MieMod.F90:857-936, CB/Chem_MieTableMod.F90:612-627).
  An off-by-one index or swapped weights returns a neighbour node instead.
* RH-linear tables: values linear in RH are reproduced at the snapped RH
  s = (int((min(rh, 0.99) + 0.001) 1000) - 1) / 1000 (slot 1 below 0) to
  float32 rounding, for any RH; a wrong snap rule moves s by 0.001 and the
  value by 1 (bound 2e-3).
* one tracer: getAOP's sums reduce to that tracer's own tau, ssa, g and
  moments (the normalisations divide back out, within 2 ulp);
* two tracers: tau adds; ssa, g and the moments are the scattering-weighted
  means (omega tau and omega tau weights) of the tracers' values, evaluated
  here in float64 (GEO/VLIDORT_MieMod.F90:283-334).
"""
import numpy as np

from aeroapps_amd import synthetic as S

f32 = np.float32
NRH = 36
NMOM, NPOL = 8, 1


def node_table(name="XX", tracers=("xx001",), seed=5, linear=False):
    """One species table on RH_i = float32(20 i) / 1000 (i = 0..35).
    linear=True: every field is a + b RH_i (node values rounded to float32)."""
    rh = (np.arange(NRH, dtype=np.float32) * f32(20.0) / f32(1000.0)).astype(np.float32)
    nb = len(tracers)
    u = lambda f: S.uniform(seed, f, np.arange(nb)[:, None, None], np.zeros((1, 1, 1), np.int64),
                            np.arange(NRH)[None, None, :])
    if linear:
        r = rh.astype(np.float64)[None, None, :]
        bext = (100.0 + 1000.0 * r + np.zeros((nb, 1, 1))).astype(np.float32)
        bsca = (80.0 + 900.0 * r + np.zeros((nb, 1, 1))).astype(np.float32)
        g = (0.5 + 0.25 * r + np.zeros((nb, 1, 1))).astype(np.float32)
    else:
        bext = (10.0 ** (2.0 + 2.0 * u(31))).astype(np.float32)
        bsca = (bext * (0.8 + 0.2 * u(32))).astype(np.float32)
        g = (0.5 + 0.3 * u(33)).astype(np.float32)
    ell = np.arange(NMOM, dtype=np.float64)[:, None, None, None]
    pm = ((2 * ell + 1) * g.astype(np.float64)[None] ** ell)[None].astype(np.float32)   # [1, nmom, nb, 1, nrh]
    return S.MieTable(name=name, tracers=tuple(tracers), rh=rh, lambda_m=np.array([f32(550e-9)], np.float32),
                      bext=bext, bsca=bsca, bbck=(bext * f32(0.05)).astype(np.float32),
                      g=g, refr=np.full_like(bext, 1.5), refi=np.full_like(bext, 0.01),
                      reff=np.full((nb, NRH), 1e-7, np.float32), pmom=pm)


def snapped_rh(rh):
    rh = np.asarray(rh, dtype=np.float32)
    isnap = np.trunc(((np.minimum(rh, f32(0.99)) + f32(0.001)).astype(f32) * f32(1000.0)).astype(f32))
    isnap = np.maximum(isnap.astype(np.int64), 1)
    return (isnap - 1) / 1000.0


def check_node_values(query):
    """query(table, q, rh) -> dict(tau, ssa, g, pmom[n, nmom]) for bin 0, channel 0."""
    t = node_table()
    i = np.arange(NRH - 1)            # the last node: slot 700 + ... beyond the grid top handled below
    rh = (t.rh[i].astype(np.float64) + 0.0005).astype(np.float32)
    q = np.full(len(i), f32(3.0e-6), np.float32)
    out = query(t, q, rh)
    assert np.array_equal(out["tau"], (t.bext[0, 0, i] * q).astype(f32))
    assert np.array_equal(out["ssa"], (t.bsca[0, 0, i] / t.bext[0, 0, i]).astype(f32))
    assert np.array_equal(out["g"], t.g[0, 0, i])
    assert np.array_equal(out["pmom"], t.pmom[0, :, 0, 0, i])
    # above the last node (0.70) every RH up to 0.99 returns the top node
    rh_top = np.array([0.7005, 0.8, 0.95, 0.99, 1.2], np.float32)
    out = query(t, np.full(5, f32(1e-6)), rh_top)
    assert np.array_equal(out["g"], np.full(5, t.g[0, 0, -1]))


def check_linear_table(query):
    t = node_table(linear=True)
    rng = np.random.default_rng(3)
    rh = np.concatenate([rng.uniform(-0.2, 0.69, 4000), [-1.0, 0.0, 0.0005, 0.3405, 0.69]]).astype(np.float32)
    rh = rh[snapped_rh(rh) <= 0.70]                 # inside the grid
    out = query(t, np.ones(len(rh), np.float32), rh)
    s = snapped_rh(rh)
    assert np.max(np.abs(out["tau"] - (100.0 + 1000.0 * s))) < 2e-3
    assert np.max(np.abs(out["g"] - (0.5 + 0.25 * s))) < 1e-6
    assert np.max(np.abs(out["ssa"] - (80.0 + 900.0 * s) / (100.0 + 1000.0 * s))) < 1e-6


def check_single_tracer(get_aop):
    """get_aop(tables, vnames, qm[nobs, nq, km], rh[nobs, km]) -> tau, ssa, g [nobs, nch, km],
    pmom[npol, nmom, nobs, nch, km]."""
    t = node_table(tracers=("xx001",))
    rng = np.random.default_rng(7)
    km, nobs = 9, 5
    qm = rng.uniform(1e-7, 1e-5, (nobs, 1, km)).astype(np.float32)
    node = rng.integers(0, NRH - 1, (nobs, km))
    rh = (t.rh[node].astype(np.float64) + 0.0005).astype(np.float32)
    tau, ssa, g, pm = get_aop([t], ["xx001"], qm, rh)
    # the lone tracer's node values: the sums and normalisations divide back out
    assert np.array_equal(tau[:, 0], (t.bext[0, 0][node] * qm[:, 0, :]).astype(f32))
    w = (t.bsca[0, 0][node] / t.bext[0, 0][node]).astype(np.float64)
    assert np.max(np.abs(ssa[:, 0] - w)) < 2e-7
    assert np.max(np.abs(g[:, 0] - t.g[0, 0][node])) < 2e-7
    for m in range(NMOM):
        pe = t.pmom[0, m, 0, 0][node].astype(np.float64)
        assert np.max(np.abs(pm[0, m, :, 0] - pe) / np.maximum(np.abs(pe), 1e-3)) < 4e-7, m


def check_two_tracers(get_aop):
    t = node_table(tracers=("xx001", "xx002"), seed=9)
    rng = np.random.default_rng(11)
    km, nobs = 7, 4
    qm = rng.uniform(1e-7, 1e-5, (nobs, 2, km)).astype(np.float32)
    # RH just above a node: each tracer contributes exactly its node values
    node = rng.integers(0, NRH - 1, (nobs, km))
    rh = (t.rh[node].astype(np.float64) + 0.0005).astype(np.float32)
    tau, ssa, g, pm = get_aop([t], ["xx001", "xx002"], qm, rh)
    b = t.bext[:, 0][:, node].astype(np.float64)       # [2, nobs, km]
    w = (t.bsca[:, 0][:, node] / t.bext[:, 0][:, node]).astype(np.float64)
    gg = t.g[:, 0][:, node].astype(np.float64)
    ti = b * np.moveaxis(qm, 1, 0).astype(np.float64)
    T = ti.sum(0)
    assert np.max(np.abs(tau[:, 0] / T - 1)) < 1e-6
    W = (w * ti).sum(0) / T
    assert np.max(np.abs(ssa[:, 0] - W)) < 2e-6
    G = (gg * w * ti).sum(0) / (W * T)
    assert np.max(np.abs(g[:, 0] - G)) < 2e-6
    for m in range(NMOM):
        P = (t.pmom[0, m][:, 0][:, node].astype(np.float64) * w * ti).sum(0) / (W * T)
        assert np.max(np.abs(pm[0, m, :, 0] - P) / np.maximum(np.abs(P), 1e-3)) < 1e-5, m
