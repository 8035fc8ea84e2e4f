"""Inputs of the sampled golden cases, regenerated from the seeded generator.

The sampled fixtures (tests/golden/sampled/*.npz) hold only the reference's
outputs, the Fourier-term counts and the case specification; their inputs are
rebuilt here from aeroapps_amd.synthetic (counter-based, so a column's values
depend only on (seed, column index)) and the float32 Chem_Mie restatement
oracle.chem_mie, and a SHA-256 over the input arrays stored with each fixture
guards against any drift of either.

Case kinds:
  lambert    scalar or vector Lambertian driver, one (km, nstreams) per case
  uneven     per-column layer counts 30..72 and stream counts {2,4,6,8}: one
             reference driver call per column with its own km and nstreams
  high_sza   as lambert with the sun at 75..89.5 degrees, views up to 70
             degrees and relative azimuth 0 or 180 degrees
  grazing    the sun at 88..89.9 degrees, views 50..80 degrees, relative
             azimuth 0 degrees for three columns in four (180 for the rest):
             points high on the outgoing line of sight then see the sun below
             the horizon, so the first-order geometry takes the sun paths
             through a tangent point
  multigeom  vector MultiGeom driver: two geometries per column
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from aeroapps_amd import synthetic as S  # noqa: E402
from oracle import chem_mie as CM  # noqa: E402

SAMPLED_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sampled")

# field ids of the overrides (disjoint from aeroapps_amd.synthetic's)
_F_HSZA, _F_HVZA, _F_HRAA, _F_G2SZA, _F_G2VZA, _F_G2RAA, _F_G2ALB = range(60, 67)
_F_GSZA, _F_GVZA, _F_GRAA = range(67, 70)

SAMPLED = {
    # the CPU-runnable workload, all of it: 1,024 columns, 550 nm, 8 streams
    "cfg0_scalar_1024": dict(kind="lambert", ncol=1024, col0=0, seed=7, channels=(550.0,), nstreams=4,
                             nmom=32, vector=False),
    # 256 columns of the 500,000-column, 4-channel, 12-stream workload
    "cfg1_scalar_256x4": dict(kind="lambert", ncol=256, col0=250_000, seed=7,
                              channels=(354.0, 388.0, 550.0, 870.0), nstreams=6, nmom=64, vector=False),
    # 64 columns of the vector 200,000-column workload
    "cfg2_vector_64x2": dict(kind="lambert", ncol=64, col0=100_000, seed=7, channels=(550.0, 870.0),
                             nstreams=4, nmom=32, vector=True),
    # 64 columns of the uneven 300,000-column workload
    "cfg3_uneven_64": dict(kind="uneven", ncol=64, col0=150_000, seed=7, channels=(550.0,), nmom=32,
                           vector=False),
    "high_sza_scalar": dict(kind="high_sza", ncol=48, col0=900, seed=11, channels=(550.0, 870.0), nstreams=6,
                            nmom=64, vector=False),
    "high_sza_vector": dict(kind="high_sza", ncol=16, col0=950, seed=11, channels=(550.0,), nstreams=4,
                            nmom=32, vector=True),
    "grazing_scalar": dict(kind="grazing", ncol=32, col0=1200, seed=13, channels=(550.0, 870.0), nstreams=6,
                           nmom=64, vector=False),
    "grazing_vector": dict(kind="grazing", ncol=16, col0=1300, seed=13, channels=(550.0,), nstreams=4,
                           nmom=32, vector=True),
    "vector_multigeom": dict(kind="multigeom", ncol=12, col0=40, seed=5, channels=(550.0, 870.0), nstreams=4,
                             nmom=32, vector=True, ngeom=2),
    "vector_2os": dict(kind="lambert", ncol=12, col0=70, seed=5, channels=(550.0, 870.0), nstreams=6,
                       nmom=32, vector=True, do_2os=True),
    # 48 columns of the vector workload over the polarised MODIS land surface
    "bpdf_modis_48x2": dict(kind="bpdf_modis", ncol=48, col0=120_000, seed=7, channels=(550.0, 870.0),
                            nstreams=4, nmom=32, vector=True),
}


def spec_json(spec: dict) -> str:
    return json.dumps(spec, sort_keys=True)


def sampled_inputs(spec: dict) -> dict:
    """Inputs of one sampled case in the reversed-Fortran C order of the oracle
    bindings (see oracle.LambertDriver)."""
    ncol, ch, nmom, vector = spec["ncol"], tuple(spec["channels"]), spec["nmom"], spec["vector"]
    npol = 6 if vector else 1
    kind = spec["kind"]
    uneven = (30, 72) if kind == "uneven" else None
    cols = S.make_columns(ncol, 72, ch, seed=spec["seed"], col0=spec["col0"], uneven_layers=uneven)
    tabs = S.make_mie_tables(ch, nmom=nmom, npol=npol)
    tau, ssa, g, pm, rc = CM.get_aop_vector(tabs, ch, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, nmom, npol)
    assert rc == 0
    sza, vza, raa = cols.sza.copy(), cols.vza.copy(), cols.raa.copy()
    idx = np.arange(spec["col0"], spec["col0"] + ncol, dtype=np.int64)
    if kind == "high_sza":
        sza = 75.0 + 14.5 * S.uniform(spec["seed"], _F_HSZA, idx)
        vza = 70.0 * S.uniform(spec["seed"], _F_HVZA, idx)
        raa = np.where(S.uniform(spec["seed"], _F_HRAA, idx) < 0.5, 0.0, 180.0)
    if kind == "grazing":
        sza = 88.0 + 1.9 * S.uniform(spec["seed"], _F_GSZA, idx)
        vza = 50.0 + 30.0 * S.uniform(spec["seed"], _F_GVZA, idx)
        raa = np.where(S.uniform(spec["seed"], _F_GRAA, idx) < 0.75, 0.0, 180.0)
    d = dict(km=72, channels=np.asarray(ch, float), nmom=nmom, npol=npol, tau=tau, ssa=ssa, g=g, pmom=pm,
             rot=cols.rot, depol=cols.depol, pe=cols.pe, he=cols.he, te=cols.te, albedo=cols.albedo.copy(),
             sza=sza, vza=vza, raa=raa, flux_factor=cols.flux_factor, qm=cols.qm, rh=cols.rh)
    if kind == "bpdf_modis":
        d["kernel_wt"], d["param"] = S.make_modis_kernels(ncol, len(ch), seed=spec["seed"], col0=spec["col0"])
        d["bpdf"] = S.make_bpdf(ncol, len(ch), seed=spec["seed"], col0=spec["col0"])
    if kind == "uneven":
        d["nlayers"] = cols.nlayers.astype(np.int32)
        d["nstreams_col"] = S.stream_choice(ncol, seed=spec["seed"], col0=spec["col0"]).astype(np.int32)
    if kind == "multigeom":
        ng = spec["ngeom"]
        assert ng == 2
        # geometry 1: the column's own; geometry 2: a second seeded view of the same column
        sza2 = 70.0 * S.uniform(spec["seed"], _F_G2SZA, idx)
        vza2 = 60.0 * S.uniform(spec["seed"], _F_G2VZA, idx)
        raa2 = 180.0 * S.uniform(spec["seed"], _F_G2RAA, idx)
        sza2[5] = 0.0    # overhead sun in geometry 2 of column 5: nadir first-order mode from here on
        d["sza"] = np.stack([sza, sza2])
        d["vza"] = np.stack([vza, vza2])
        d["raa"] = np.stack([raa, raa2])
        alb2 = 0.02 + 0.1 * S.uniform(spec["seed"], _F_G2ALB, idx)
        d["albedo"] = np.stack([d["albedo"], np.broadcast_to(alb2, d["albedo"].shape)])
    return d


def input_checksum(d: dict) -> str:
    h = hashlib.sha256()
    for k in sorted(d):
        v = d[k]
        if isinstance(v, np.ndarray):
            h.update(k.encode())
            h.update(np.ascontiguousarray(v).tobytes())
    return h.hexdigest()


def load_sampled(name: str) -> dict:
    z = np.load(os.path.join(SAMPLED_DIR, name + ".npz"), allow_pickle=False)
    out = {k: z[k] for k in z.files}
    out["spec"] = json.loads(str(out["spec"]))
    return out


def sampled_names():
    if not os.path.isdir(SAMPLED_DIR):
        return []
    return sorted(f[:-4] for f in os.listdir(SAMPLED_DIR) if f.endswith(".npz"))
