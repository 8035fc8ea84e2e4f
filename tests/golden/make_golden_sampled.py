"""Writes the sampled golden fixtures (tests/golden/sampled/*.npz) from the
reference's own AeroApps drivers.

Outputs come from VLIDORT_Scalar_Lambert / VLIDORT_Vector_Lambert
(SingleGeom and MultiGeom) compiled from the reference sources
(oracle/Makefile.ref -> oracle/_ref/libvlidort_drv.so, called through
oracle/ref_calls.f90).  The drivers do not return the number of Fourier terms;
the oracle's statement-by-statement harness (oracle/ref_driver.f90) does, and
this script requires its radiances to equal the drivers' bit for bit before
it stores that count.  Where no harness run reproduces the driver's state
(MultiGeom geometries after the overhead-sun latch set by another geometry)
the count is stored as -1 and the tests do not compare it.

Inputs are not stored: tests/golden/golden_inputs.py rebuilds them from the
seeded generator and the float32 Chem_Mie restatement, and the fixture keeps
their SHA-256.  The VLIDORT_ROT_CALC fixture (prep/rot_calc.npz) stores
the reference's Rayleigh optical thickness for the pressure/height/temperature
edges of 32 generated columns at 7 wavelengths, and the depolarisation ratio
from the reference's coef_depol (VLIDORT_ROT_CALC's own depol output reads a
local of VLIDORT_Rayleigh through a pointer after that routine has returned,
so it holds whatever the stack holds).

Run here (the reference is only available in this container):
    python tests/golden/make_golden_sampled.py [case ...]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from golden_inputs import (ROOT, SAMPLED, SAMPLED_DIR, input_checksum, sampled_inputs,  # noqa: E402
                           spec_json)

sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from aeroapps_amd import synthetic as S  # noqa: E402

f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731


def lambert(D, H, d, km, N, vector, do_2os=False, sl=slice(None)):
    """Reference driver + harness on columns `sl` with km layers."""
    ch, nmom = d["channels"], d["nmom"]
    rot = f64(d["rot"][:, sl, :km])
    tau, ssa, g = (f64(d[k][sl, :, :km]) for k in ("tau", "ssa", "g"))
    pm = f64(d["pmom"][:, :, sl, :, :km])
    pe, he, te = (f64(d[k][sl, :km + 1]) for k in ("pe", "he", "te"))
    alb = f64(d["albedo"][:, sl])
    sza, raa, vza = (f64(d[k][sl]) for k in ("sza", "raa", "vza"))
    if vector:
        alpha = np.zeros_like(rot)
        ff = f64(d["flux_factor"][sl])
        I, R, Q, U, rc = D.vector_lambert(km, ch, N, nmom, rot, d["depol"], alpha, tau, ssa, pm, pe, he, te, alb,
                                          sza, raa, vza, ff, do_2os=do_2os)
        if do_2os:
            # NSTOKES = 1 through the vector driver: the scalar harness computes
            # the same radiance (same Greek assembly for entry 1, alpha = 0, F0 = 1)
            I1, R1, nf, rc1 = H.scalar_lambert(km, ch, N, nmom, rot, d["depol"], tau, ssa, g, pm[:1], pe, he, te,
                                               alb, sza, raa, vza)
            assert np.array_equal(I, I1) and np.array_equal(R, R1) and rc == rc1
        else:
            I1, R1, Q1, U1, nf, rc1 = H.vector_lambert(km, ch, N, nmom, rot, d["depol"], alpha, tau, ssa, pm, pe, he,
                                                       te, alb, sza, raa, vza, ff)
            assert all(np.array_equal(a, b) for a, b in ((I, I1), (R, R1), (Q, Q1), (U, U1))) and rc == rc1
        return dict(radiance=I, reflectance=R, q=Q, u=U, nfourier=nf, rc=np.int32(rc))
    I, R, rc = D.scalar_lambert(km, ch, N, nmom, rot, d["depol"], tau, ssa, g, pm, pe, he, te, alb, sza, raa, vza)
    I1, R1, nf, rc1 = H.scalar_lambert(km, ch, N, nmom, rot, d["depol"], tau, ssa, g, pm, pe, he, te, alb, sza, raa,
                                       vza)
    assert np.array_equal(I, I1) and np.array_equal(R, R1) and rc == rc1
    return dict(radiance=I, reflectance=R, nfourier=nf, rc=np.int32(rc))


def run_case(D, H, spec, d):
    kind = spec["kind"]
    if kind in ("lambert", "high_sza", "grazing"):
        return lambert(D, H, d, 72, spec["nstreams"], spec["vector"], spec.get("do_2os", False))
    if kind == "bpdf_modis":
        M = oracle.drv_harness_modis()
        common = (72, d["channels"], spec["nstreams"], spec["nmom"], d["rot"], d["depol"], f64(d["tau"]),
                  f64(d["ssa"]), f64(d["pmom"]), d["pe"], d["he"], d["te"])
        rad, refl, q, u, br, brq, bru, nf, rc = M.vector_bpdf(*common, 1, d["albedo"], d["kernel_wt"], d["param"],
                                                              d["bpdf"], d["sza"], d["raa"], d["vza"])
        want = D.vector_landmodis_bpdf(*common, d["kernel_wt"], d["param"], d["bpdf"], d["sza"], d["raa"],
                                       d["vza"])
        for a, b in zip((rad, refl, q, u, br, brq, bru), want[:7]):
            assert np.array_equal(a, b)
        assert rc == want[-1]
        return dict(radiance=rad, reflectance=refl, q=q, u=u, br=br, br_q=brq, br_u=bru, nfourier=nf,
                    rc=np.int32(rc))
    if kind == "uneven":
        nch, n = len(spec["channels"]), spec["ncol"]
        out = dict(radiance=np.zeros((nch, n)), reflectance=np.zeros((nch, n)),
                   nfourier=np.zeros((nch, n), np.int32), rc=np.int32(0))
        for o in range(n):
            r = lambert(D, H, d, int(d["nlayers"][o]), int(d["nstreams_col"][o]), False, sl=slice(o, o + 1))
            assert r["rc"] == 0
            for k in ("radiance", "reflectance", "nfourier"):
                out[k][:, o] = r[k][:, 0]
        return out
    if kind == "multigeom":
        ch, nmom, N = d["channels"], d["nmom"], spec["nstreams"]
        alpha = np.zeros_like(d["rot"])
        args = (72, ch, N, nmom, d["rot"], d["depol"], alpha, f64(d["tau"]), f64(d["ssa"]), f64(d["pmom"]),
                d["pe"], d["he"], d["te"])
        I, R, Q, U, rc = D.vector_lambert_mg(*args, d["albedo"], d["sza"], d["raa"], d["vza"], d["flux_factor"])
        ng, nch, n = I.shape
        nf = np.full((ng, nch, n), -1, np.int32)
        # the overhead-sun latch of any geometry switches every geometry of that
        # column and of the later columns to nadir first-order mode
        zero = np.nonzero((d["sza"] == 0.0).any(axis=0))[0]
        jl = int(zero[0]) if len(zero) else n
        for g in range(ng):
            I1, R1, Q1, U1, nf1, rc1 = H.vector_lambert(*args, d["albedo"][0], d["sza"][g], d["raa"][g], d["vza"][g],
                                                        d["flux_factor"])
            own = (d["sza"][g] == 0.0).any() and (d["sza"][g][jl] == 0.0 if jl < n else True)
            upto = n if own else jl
            for a, b in ((I[g], I1), (R[g], R1), (Q[g], Q1), (U[g], U1)):
                assert np.array_equal(a[:, :upto], b[:, :upto])
            nf[g][:, :upto] = nf1[:, :upto]
        return dict(radiance=I, reflectance=R, q=Q, u=U, nfourier=nf, rc=np.int32(rc))
    raise ValueError(kind)


def rot_fixture(D):
    ch = np.array([340.0, 354.0, 388.0, 470.0, 550.0, 660.0, 870.0])
    cols = S.make_columns(32, 72, (550.0,), seed=3, col0=77)
    rot, _, rc = D.rot_calc(72, ch, cols.pe, cols.he, cols.te)
    assert rc == 0
    # the wavelength in micron exactly as VLIDORT_Rayleigh forms it (default-real 1.E-3)
    depol = np.array([D.coef_depol(c * float(np.float32(1e-3))) for c in ch])
    os.makedirs(os.path.join(HERE, "prep"), exist_ok=True)
    np.savez_compressed(os.path.join(HERE, "prep", "rot_calc.npz"), channels=ch, pe=cols.pe, he=cols.he, te=cols.te,
                        out_rot=rot, out_depol=depol)
    print("rot_calc", rot.shape, depol)


def main():
    D, H = oracle.drivers(), oracle.drv_harness()
    if D is None:
        raise SystemExit("reference driver library not built: make -C oracle -f Makefile.ref")
    os.makedirs(SAMPLED_DIR, exist_ok=True)
    only = set(sys.argv[1:])
    if not only or "rot_calc" in only:
        rot_fixture(D)
    for name, spec in SAMPLED.items():
        if only and name not in only:
            continue
        t = time.time()
        d = sampled_inputs(spec)
        out = run_case(D, H, spec, d)
        np.savez_compressed(os.path.join(SAMPLED_DIR, name + ".npz"), spec=np.str_(spec_json(spec)),
                            checksum=np.str_(input_checksum(d)), **{"out_" + k: v for k, v in out.items()})
        print(name, "rc", int(out["rc"]), "nfourier", np.unique(out["nfourier"]), f"{time.time() - t:.1f} s")


if __name__ == "__main__":
    main()
