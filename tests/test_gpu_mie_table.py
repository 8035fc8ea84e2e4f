"""Tables built by the library from optics-file arrays
(aa_mie_add_table_native: Chem_MieTableRead's fills, derived fields, polint)
and the extra Chem_MieQuery outputs equal the oracle restatement bit for bit;
the optics of getAOPvector from such a table equal those from the oracle's
interpolated table added pre-interpolated."""
import numpy as np
import pytest

from aeroapps_amd import chem_mie as AM
from aeroapps_amd import synthetic as S
from oracle import chem_mie as CM
from test_mie_table import _file_arrays

pytestmark = pytest.mark.gpu
f32 = np.float32


def test_native_table_matches_oracle(gpu):
    print("case native table", flush=True)
    rng = np.random.default_rng(21)
    d = _file_arrays(rng, nl=3, nrh=5, nbin=2, nmom=8, npol=6)
    ch = np.array([4.5e-7, 5.5e-7, 7.0e-7], dtype=f32)   # below, inside and above the file's wavelengths
    rc0, t = CM.table_read(d["lambda_table"], ch, d["rh"], d["reff"], d["bext"], d["bsca"], d["bbck"], d["g"],
                           nmom=5, pback=d["pback"], pmom=d["pmom"], gf=d["gf"], rhop=d["rhop"], refr=d["refr"],
                           refi=d["refi"])
    T = lambda a: None if a is None else np.asarray(a).T          # reversed-C -> Fortran shapes
    mie = AM.MieTables(ch)
    rc = mie.add_table_native(d["lambda_table"], d["rh"], T(d["reff"]), T(d["bext"]), T(d["bsca"]), T(d["bbck"]),
                              T(d["g"]), ["du001", "du002"], nmom=5, pback=T(d["pback"]), pmom=T(d["pmom"]),
                              gf=T(d["gf"]), rhop=T(d["rhop"]), refr=T(d["refr"]), refi=T(d["refi"]))
    assert rc == rc0 == 0
    rh = rng.uniform(0.0, 1.0, 64).astype(f32)
    q = rng.uniform(1e-6, 1e-4, 64).astype(f32)
    for ib in range(2):
        for c in range(3):
            got = mie.query(ib, c + 1, q, rh, nmom=5, npol=6)
            tau, ssa, g, pm = CM.query(t["bext"], t["bsca"], t["g"], t["pmom"], t["rhi"], t["rha"], ib, c, q, rh)
            assert np.array_equal(got["tau"], tau) and np.array_equal(got["ssa"], ssa)
            assert np.array_equal(got["gasym"], g) and np.array_equal(got["pmom"], pm)
            ex = mie.query_extra(ib, c + 1, rh)
            want = CM.query_extra(t, ib, c, rh)
            for k in ("reff", "p11", "p22", "gf", "rhod", "vol", "area"):
                assert np.array_equal(ex[k], want[k]), k
    # more moments than the file holds: rc 99, nothing added
    assert mie.add_table_native(d["lambda_table"], d["rh"], T(d["reff"]), T(d["bext"]), T(d["bsca"]), T(d["bbck"]),
                                T(d["g"]), ["x1", "x2"], nmom=9, pmom=T(d["pmom"])) == 99
    assert mie.query_idx("x1") == -1
    mie.close()


def test_get_aop_from_native_tables(gpu):
    """A synthetic species set written as file arrays on a coarser wavelength
    grid, read with polint, gives the same getAOPvector optics as the oracle's
    interpolated tables."""
    print("case native getAOP", flush=True)
    channels = (470.0, 550.0, 660.0)
    nmom, npol = 16, 6
    tabs = S.make_mie_tables(channels, nmom=nmom, npol=npol)
    lam_file = np.array([4.0e-7, 5.0e-7, 6.0e-7, 7.0e-7]) + 1e-17
    ch_m = (np.asarray(channels) * 1e-9).astype(f32)
    mie = AM.MieTables(ch_m)
    rng = np.random.default_rng(3)
    ref_tables = []
    for tb in tabs:
        nbin, nrh = tb.nbin, len(tb.rh)
        # file arrays: channel-3 table values spread over the file grid with a smooth slope
        fl = lambda base: np.stack([base * (1.0 + 0.1 * k) for k in range(4)], axis=-1)   # [nbin, nrh, 4]
        bext, bsca, g = fl(tb.bext[:, 1].astype(np.float64)), fl(tb.bsca[:, 1].astype(np.float64)), \
            fl(tb.g[:, 1].astype(np.float64))
        pm = np.stack([tb.pmom[:, :, :, 1].astype(np.float64) * (1.0 + 0.05 * k) for k in range(4)], axis=-1)
        reff = rng.uniform(1e-7, 1e-6, (nbin, nrh))
        rc, t = CM.table_read(lam_file, ch_m, tb.rh, reff, bext, bsca, bext * 0.1, g, nmom=nmom, pmom=pm)
        assert rc == 0
        ref_tables.append((t, tb.tracers))
        T = lambda a: np.asarray(a).T
        assert mie.add_table_native(lam_file, tb.rh, T(reff), T(bext), T(bsca), T(bext * 0.1), T(g), tb.tracers,
                                    nmom=nmom, pmom=T(pm)) == 0
    cols = S.make_columns(6, 72, channels, seed=9)
    tau, ssa, gg, pmo, rc = AM.get_aop_vector(mie, np.asarray(channels), S.TRACERS, 0, cols.qm.T, cols.rh.T, nmom,
                                              npol)
    # the oracle's getAOPvector over the interpolated tables
    otabs = []
    for (t, names), tb in zip(ref_tables, tabs):
        o = type(tb)(**{**tb.__dict__})
        o.bext, o.bsca, o.g, o.pmom = t["bext"], t["bsca"], t["g"], t["pmom"]
        otabs.append(o)
    t0, s0, g0, p0, rc0 = CM.get_aop_vector(otabs, channels, ch_m, S.TRACERS, cols.qm, cols.rh, nmom, npol)
    assert rc == rc0 == 0
    assert np.array_equal(tau, t0.T) and np.array_equal(ssa, s0.T) and np.array_equal(gg, g0.T)
    assert np.array_equal(pmo, p0.T)
    mie.close()
