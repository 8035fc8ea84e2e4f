"""Columns per second of the vector (three Stokes components) host-pointer
drivers by surface: Lambertian (VLIDORT_Vector_Lambert), Cox-Munk ocean
(VLIDORT_Vector_CX_SingleGeom), Lambertian + BPDF (VLIDORT_Vector_Lambert_BPDF)
and MODIS + BPDF (VLIDORT_Vector_LandMODIS_BPDF), on seeded synthetic columns
(optics from the device Chem_Mie query).  Each timed call follows an untimed
one; times include the host-device copies of the drivers' arrays.

    python -m aeroapps_amd.surface_rates [--columns 20000] [--surface all] [--passes 1]
"""
import argparse
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--columns", type=int, default=20000)
    ap.add_argument("--nstreams", type=int, default=4)
    ap.add_argument("--surface", default="all", choices=("all", "lambert", "cx", "lambert_bpdf", "modis_bpdf"))
    ap.add_argument("--precision", type=int, default=-1, help="0 FP64, 1 mixed, -1 both")
    ap.add_argument("--passes", type=int, default=1, help="timed calls per (surface, precision)")
    a = ap.parse_args(argv)
    import torch
    from . import chem_mie as AM
    from . import synthetic as S
    from . import vlidort as V
    ncol, ch, N, nmom = a.columns, (550.0, 870.0), a.nstreams, 32
    cols = S.make_columns(ncol, 72, ch, seed=7)
    mie = AM.MieTables.from_synthetic(S.make_mie_tables(ch, nmom=nmom, npol=6))
    tau, ssa, g, pm, rc = AM.get_aop_vector(mie, np.asarray(ch, np.float32), S.TRACERS, 0, cols.qm.T, cols.rh.T,
                                            nmom, 6)
    kw, par = S.make_modis_kernels(ncol, len(ch), seed=7)
    bp = S.make_bpdf(ncol, len(ch), seed=7)
    u, v, mr = S.make_ocean(ncol, ch, seed=7)
    rot = cols.rot.transpose(2, 1, 0)
    common = (cols.channels, N, False, rot, cols.depol)
    geo = (cols.pe.T, cols.he.T, cols.te.T)
    ang = (cols.sza, cols.raa, cols.vza, S.MISSING)
    runs = {
        "lambert": lambda p: V.vector_lambert(*common, np.zeros_like(rot), tau, ssa, pm, *geo, cols.albedo.T,
                                              *ang[:3], cols.flux_factor.T, S.MISSING, precision=p),
        "cx": lambda p: V.vector_cx(*common, tau, ssa, pm, *geo, u, v, mr, *ang, precision=p),
        "lambert_bpdf": lambda p: V.vector_lambert_bpdf(*common, tau, ssa, pm, *geo, cols.albedo.T, bp.T, *ang,
                                                        precision=p),
        "modis_bpdf": lambda p: V.vector_brdf_modis_bpdf(*common, tau, ssa, pm, *geo, kw.T, par.T, bp.T, *ang,
                                                         precision=p),
    }
    precs = (0, 1) if a.precision < 0 else (a.precision,)
    for name, fn in runs.items():
        if a.surface not in ("all", name):
            continue
        for p in precs:
            fn(p)
            for _ in range(a.passes):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                fn(p)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                print(f"{name} precision={p}: {ncol} columns x {len(ch)} channels, N={N}: {dt:.3f} s, "
                      f"{ncol / dt:.0f} columns/s", flush=True)
    mie.close()


if __name__ == "__main__":
    main()
