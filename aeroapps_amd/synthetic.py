"""Seeded synthetic inputs for the aerosol-optics + discrete-ordinate radiance path.

Every value is a pure function of (seed, field id, column index, inner index),
drawn from a counter-based SplitMix64 hash, so a column's inputs do not depend
on which batch, granule or chunk it is generated in.  The arrays produced have
exactly the dtypes and (Fortran column-major) shapes of the reference call
surfaces:

* ``VLIDORT_getAOPvector`` (src/Components/rtms/geo_vlidort/VLIDORT_MieMod.F90:124-145):
  ``qm(km,nq,nobs)`` and ``rh(km,nobs)`` as float32.
* ``VLIDORT_Scalar_Lambert`` / ``VLIDORT_Vector_Lambert``
  (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:34-87, 172-235): ``ROT(km,nobs,nch)``,
  ``pe/he/te(km+1,nobs)``, ``albedo(nobs,nch)``, angles ``(nobs)`` as float64.

Arrays here are returned as C-ordered numpy arrays whose *reversed* shape is
the Fortran shape (numpy ``qm[nobs, nq, km]`` is Fortran ``qm(km,nq,nobs)``);
that is the same memory layout, so they can be handed to the C-ABI and to the
oracle without copies.

Physical choices (the workload description leaves these to the builder):
* layer pressure thickness U[200, 2000] Pa, top edge pressure 1 Pa
  (``ptop``, src/Components/rtms/geo_vlidort/geo_vlidort_pars.F90:9);
* temperature from the 1976 standard-atmosphere troposphere law
  T = max(216.65, 288.15 (p/101325)^0.190263) at layer mid pressure, air
  density p/(R T); edge p, z, T then follow ``getEdgeVars``
  (src/Shared/GMAO_Shared/Chem_Base/MieObs_py.F90:27-83) in float32;
* tracer mixing ratio log-normal with median 1e-9 kg/kg and geometric standard
  deviation 1.5; ``qm = q * delp / grav`` with grav = 9.81
  (src/Components/rtms/geo_vlidort/geo_vlidort.F90:2124-2144, geo_vlidort_pars.F90:47);
* Rayleigh optical depth 0.097 at 550 nm spread by layer pressure thickness
  and scaled to other wavelengths by the Bodhaine cross-section ratio of
  ``VLIDORT_Rayleigh`` (src/Components/rtms/RTSI/VLIDORT2p8p3/VLIDORT_ScatMod.F90:209-245);
  depolarisation 0.03 for every channel;
* Mie tables: 36 RH nodes 0.99*i/35, bext log-uniform [1e2, 1e4] m2/kg, ssa
  U[0.8, 1], g U[0.5, 0.8], P11 moments (2l+1) g^l; for npol = 6 the other
  phase-matrix moments are scaled copies of P11 (P12 = -c12 P11, P33 = c33 P11,
  P34 = c34 P11 for l >= 2, P22 = P11, P44 = P33) with seeded c12 in [0, 0.5],
  c33 in [0.8, 1], c34 in [0, 0.2].
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Sequence, Tuple

import numpy as np

MISSING = 1.0e15

# Tracer names and order of the reference geo_vlidort driver
# (src/Components/rtms/geo_vlidort/geo_vlidort_pars.F90:29-33).
TRACERS: Tuple[str, ...] = (
    "du001", "du002", "du003", "du004", "du005",
    "ss001", "ss002", "ss003", "ss004", "ss005",
    "BCphobic", "BCphilic", "OCphobic", "OCphilic", "SO4",
)
# Species tables: (table name, tracer names in bin order)
SPECIES: Tuple[Tuple[str, Tuple[str, ...]], ...] = (
    ("DU", TRACERS[0:5]),
    ("SS", TRACERS[5:10]),
    ("BC", TRACERS[10:12]),
    ("OC", TRACERS[12:14]),
    ("SU", TRACERS[14:15]),
)
NRH = 36

# field ids for the counter-based generator
_F_DELP, _F_Q1, _F_Q2, _F_RH, _F_SZA, _F_VZA, _F_RAA, _F_NLAY, _F_NSTR = range(1, 10)
_F_FISO, _F_FVOL, _F_FGEO, _F_U10, _F_V10 = range(10, 15)
_F_NDVI, _F_BPDFC = range(15, 17)
_F_TBEXT, _F_TSSA, _F_TG, _F_TC12, _F_TC33, _F_TC34, _F_TREFR, _F_TREFI, _F_TREFF = range(20, 29)

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on a uint64 array (wrap-around arithmetic)."""
    x = x.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        x += np.uint64(0x9E3779B97F4A7C15)
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x


def uniform(seed: int, field: int, *idx: np.ndarray) -> np.ndarray:
    """Uniform float64 in [0, 1) keyed by (seed, field, idx...), broadcast over idx."""
    key = np.uint64((seed * 0x100000001B3 + field * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    h = _splitmix64(np.asarray(key, dtype=np.uint64))
    arrs = np.broadcast_arrays(*[np.asarray(i, dtype=np.int64) for i in idx])
    out = np.broadcast_to(h, arrs[0].shape).copy() if arrs else h
    for a in arrs:
        out = _splitmix64(out ^ a.astype(np.uint64))
    return (out >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _rayleigh_xsec(wmicron: float) -> float:
    """Bodhaine/Tomasi molecular cross section A(lambda) as in
    VLIDORT_ScatMod.F90:209-245 (only its ratio between channels is used)."""
    xx = 1.0 / wmicron
    xx2 = xx * xx
    cm = wmicron / 10000.0
    ri = 8060.51 + 2480990.0 / (132.274 - xx2) + 17455.7 / (39.32957 - xx2)
    ri = ri * 1e-8 + 1.0
    c = 360.0e-6
    ri = (ri - 1.0) * (1.0 + 0.54 * (c - 0.0003)) + 1.0
    ri2 = ri * ri
    av = 6.0221367e23
    ts = 288.15
    ns = (av / 22.4141) * (273.15 / ts) * (1.0 / 1000)
    d1 = 1.034 + 3.17e-4 * xx2
    d2 = 1.096 + 1.385e-3 * xx2 + 1.448e-4 * xx2 * xx2
    fair = (78.084 * d1 + 20.946 * d2 + 0.934 * 1.0 + c * 100.0 * 1.15) / (78.084 + 20.946 + 0.934 + c * 100.0)
    pi3 = math.pi ** 3
    return 24.0 * pi3 * fair * ((ri2 - 1) ** 2) / (ns * ns * cm ** 4 * (ri2 + 2) ** 2)


@dataclasses.dataclass
class MieTable:
    """One species table with the fields of ``Chem_MieTable``
    (src/Shared/GMAO_Shared/Chem_Base/Chem_MieTableMod.F90:55-83), already on the
    requested channels.  Arrays are C-ordered with reversed Fortran shape:
    ``bext[nbin, nch, nrh]`` is Fortran ``bext(nrh,nch,nbin)``;
    ``pmom[npol, nmom, nbin, nch, nrh]`` is Fortran ``pmom(nrh,nch,nbin,nmom,npol)``."""
    name: str
    tracers: Tuple[str, ...]
    rh: np.ndarray          # (nrh,) float32
    lambda_m: np.ndarray    # (nch,) float32, wavelengths in metres
    bext: np.ndarray
    bsca: np.ndarray
    bbck: np.ndarray
    g: np.ndarray
    refr: np.ndarray
    refi: np.ndarray
    reff: np.ndarray        # [nbin, nrh]
    pmom: np.ndarray

    @property
    def nbin(self) -> int:
        return len(self.tracers)

    @property
    def nmom(self) -> int:
        return self.pmom.shape[1]

    @property
    def npol(self) -> int:
        return self.pmom.shape[0]


def make_mie_tables(channels_nm: Sequence[float], nmom: int, npol: int, seed: int = 2024) -> List[MieTable]:
    """Synthetic Mie tables for the 15 tracers (5 species tables)."""
    nch = len(channels_nm)
    rh = (0.99 * np.arange(NRH, dtype=np.float64) / (NRH - 1)).astype(np.float32)
    lam = (np.asarray(channels_nm, dtype=np.float64) * 1e-9).astype(np.float32)
    tables = []
    for ti, (name, tr) in enumerate(SPECIES):
        nb = len(tr)
        b = np.arange(nb)[:, None, None]
        c = np.arange(nch)[None, :, None]
        r = np.arange(NRH)[None, None, :]
        key = ti * 100 + b
        bext = (10.0 ** (2.0 + 2.0 * uniform(seed, _F_TBEXT, key, c, r))).astype(np.float32)
        ssa = 0.8 + 0.2 * uniform(seed, _F_TSSA, key, c, r)
        bsca = (bext.astype(np.float64) * ssa).astype(np.float32)
        gg = 0.5 + 0.3 * uniform(seed, _F_TG, key, c, r)
        g32 = gg.astype(np.float32)
        bbck = (bext * np.float32(0.05)).astype(np.float32)
        refr = (1.4 + 0.2 * uniform(seed, _F_TREFR, key, c, r)).astype(np.float32)
        refi = (0.001 + 0.01 * uniform(seed, _F_TREFI, key, c, r)).astype(np.float32)
        reff = (1e-7 * (1.0 + 10.0 * uniform(seed, _F_TREFF, ti * 100 + np.arange(nb)[:, None],
                                               np.arange(NRH)[None, :]))).astype(np.float32)
        ell = np.arange(nmom, dtype=np.float64)
        g64 = g32.astype(np.float64)[None, ...]                 # [1, nb, nch, nrh]
        p11 = (2.0 * ell[:, None, None, None] + 1.0) * g64 ** ell[:, None, None, None]  # [nmom, nb, nch, nrh]
        pm = np.zeros((npol, nmom, nb, nch, NRH), dtype=np.float64)
        pm[0] = p11
        if npol >= 4:
            c12 = 0.5 * uniform(seed, _F_TC12, key, c, r)[None]
            c33 = 0.8 + 0.2 * uniform(seed, _F_TC33, key, c, r)[None]
            c34 = 0.2 * uniform(seed, _F_TC34, key, c, r)[None]
            lge2 = (ell >= 2)[:, None, None, None]
            pm[1] = -c12 * p11 * lge2
            pm[2] = c33 * p11
            pm[3] = c34 * p11 * lge2
            if npol >= 6:
                pm[4] = p11
                pm[5] = pm[2]
        tables.append(MieTable(name=name, tracers=tuple(tr), rh=rh, lambda_m=lam,
                               bext=bext, bsca=bsca, bbck=bbck, g=g32, refr=refr, refi=refi,
                               reff=reff, pmom=pm.astype(np.float32)))
    return tables


@dataclasses.dataclass
class Columns:
    """A batch of synthetic columns in reference (Fortran column-major) layout."""
    km: int
    channels: np.ndarray      # (nch,) float64 [nm]
    qm: np.ndarray            # [nobs, nq, km] float32
    rh: np.ndarray            # [nobs, km] float32
    rot: np.ndarray           # [nch, nobs, km] float64
    depol: np.ndarray         # (nch,) float64
    pe: np.ndarray            # [nobs, km+1] float64
    he: np.ndarray            # [nobs, km+1] float64
    te: np.ndarray            # [nobs, km+1] float64
    albedo: np.ndarray        # [nch, nobs] float64
    sza: np.ndarray           # (nobs,) float64
    vza: np.ndarray
    raa: np.ndarray
    flux_factor: np.ndarray   # [nobs, nch] float64
    nlayers: np.ndarray       # (nobs,) int32, used layers (<= km) for uneven batches
    col0: int = 0

    @property
    def nobs(self) -> int:
        return self.qm.shape[0]

    @property
    def nch(self) -> int:
        return len(self.channels)


def _edge_vars(delp: np.ndarray, ptop: float = 1.0):
    """pe, ze, te from delp and a standard-atmosphere air density, following
    getEdgeVars (src/Shared/GMAO_Shared/Chem_Base/MieObs_py.F90:27-83) in float32."""
    f32 = np.float32
    nobs, km = delp.shape
    pe = np.empty((nobs, km + 1), dtype=f32)
    pe[:, 0] = f32(ptop)
    for k in range(km):
        pe[:, k + 1] = pe[:, k] + delp[:, k]
    pm_true = 0.5 * (pe[:, :-1].astype(np.float64) + pe[:, 1:].astype(np.float64))
    tstd = np.maximum(216.65, 288.15 * (pm_true / 101325.0) ** 0.190263)
    rgas = f32(287.0)
    grav = f32(9.80616)
    kappa = f32(2.0) / f32(7.0)
    airdens = (pm_true / (287.0 * tstd)).astype(f32)
    ze = np.empty((nobs, km + 1), dtype=f32)
    ze[:, km] = f32(0.0)
    for k in range(km - 1, -1, -1):
        ze[:, k] = ze[:, k + 1] + delp[:, k] / (airdens[:, k] * grav)
    pm = (pe[:, :-1] + pe[:, 1:]) / f32(2.0)
    tm = pm / (airdens * rgas)
    te = np.empty((nobs, km + 1), dtype=f32)
    te[:, 0] = tm[:, 0]
    for k in range(1, km):
        alpha = np.log(pe[:, k] / pm[:, k - 1]) / np.log(pm[:, k] / pm[:, k - 1])
        te[:, k] = tm[:, k - 1] + alpha * (tm[:, k] - tm[:, k - 1])
    te[:, km] = tm[:, km - 1] * (pe[:, km] / pm[:, km - 1]) ** kappa
    return pe.astype(np.float64), ze.astype(np.float64), te.astype(np.float64)


def make_columns(ncol: int, km: int, channels_nm: Sequence[float], seed: int = 7,
                 col0: int = 0, uneven_layers: Tuple[int, int] | None = None) -> Columns:
    """Columns col0 .. col0+ncol-1 of the seeded workload.

    ``uneven_layers=(lmin, lmax)`` draws a per-column layer count uniformly from
    [lmin, lmax] (terrain truncation); the arrays are still dimensioned ``km``
    and layers beyond a column's count are filled with the values of a valid
    profile, so each column is also a complete km-layer column.
    """
    cols = np.arange(col0, col0 + ncol, dtype=np.int64)
    c2 = cols[:, None]
    kk = np.arange(km)[None, :]
    nq = len(TRACERS)
    delp = (200.0 + 1800.0 * uniform(seed, _F_DELP, c2, kk)).astype(np.float32)
    c3, q3, k3 = cols[:, None, None], np.arange(nq)[None, :, None], np.arange(km)[None, None, :]
    u1 = uniform(seed, _F_Q1, c3, q3, k3)
    u2 = uniform(seed, _F_Q2, c3, q3, k3)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)
    q = 1.0e-9 * np.exp(math.log(1.5) * z)
    qm = (q.astype(np.float32) * delp[:, None, :] / np.float32(9.81)).astype(np.float32)
    rh = (0.99 * uniform(seed, _F_RH, c2, kk)).astype(np.float32)
    pe, he, te = _edge_vars(delp)
    sza = 70.0 * uniform(seed, _F_SZA, cols)
    vza = 60.0 * uniform(seed, _F_VZA, cols)
    raa = 180.0 * uniform(seed, _F_RAA, cols)
    ch = np.asarray(channels_nm, dtype=np.float64)
    dp64 = delp.astype(np.float64)
    frac = dp64 / dp64.sum(axis=1, keepdims=True)
    a550 = _rayleigh_xsec(0.55)
    rot = np.stack([0.097 * frac * (_rayleigh_xsec(w * 1e-3) / a550) for w in ch], axis=0)
    nch = len(ch)
    if uneven_layers is None:
        nl = np.full(ncol, km, dtype=np.int32)
    else:
        lo, hi = uneven_layers
        nl = (lo + np.floor((hi - lo + 1) * uniform(seed, _F_NLAY, cols))).astype(np.int32)
    return Columns(km=km, channels=ch, qm=qm, rh=rh, rot=rot,
                   depol=np.full(nch, 0.03), pe=pe, he=he, te=te,
                   albedo=np.full((nch, ncol), 0.05), sza=sza, vza=vza, raa=raa,
                   flux_factor=np.ones((ncol, nch)), nlayers=nl, col0=col0)


def stream_choice(ncol: int, choices: Sequence[int] = (2, 4, 6, 8), seed: int = 7, col0: int = 0) -> np.ndarray:
    """Per-column half-space stream count drawn from ``choices`` (uneven batches)."""
    cols = np.arange(col0, col0 + ncol, dtype=np.int64)
    i = np.floor(len(choices) * uniform(seed, _F_NSTR, cols)).astype(np.int64)
    return np.asarray(choices, dtype=np.int32)[i]


def make_modis_kernels(ncol: int, nch: int, seed: int = 7, col0: int = 0):
    """MODIS RTLS kernel weights and Li-sparse parameters per (column, channel):
    kernel_wt [ncol, nch, 3] = (fiso, fgeo, fvol) in the order of the reference's
    kernel_wt(3,nch,nobs), param [ncol, nch, 2] = (h/b, b/r) = (2, 1), the
    MODIS operational values.  Weights span typical MCD43 land values."""
    cols = np.arange(col0, col0 + ncol, dtype=np.int64)[:, None]
    chs = np.arange(nch, dtype=np.int64)[None, :]
    fiso = 0.03 + 0.27 * uniform(seed, _F_FISO, cols, chs)
    fvol = fiso * (0.1 + 0.5 * uniform(seed, _F_FVOL, cols, chs))
    fgeo = fiso * (0.02 + 0.18 * uniform(seed, _F_FGEO, cols, chs))
    kw = np.stack([fiso, fgeo, fvol], axis=-1)
    par = np.empty((ncol, nch, 2))
    par[..., 0] = 2.0
    par[..., 1] = 1.0
    return kw, par


def make_bpdf(ncol: int, nch: int, seed: int = 7, col0: int = 0):
    """BPDF-NDVI parameters per (column, channel), bpdf [ncol, nch, 3] in the
    order of the reference's BPDFparam(3,nch,nobs): refractive index 1.5, NDVI
    uniform on [-0.2, 0.9] and the scaling factor C uniform on [4, 8] (the
    range of Maignan et al. 2009, Table 1)."""
    cols = np.arange(col0, col0 + ncol, dtype=np.int64)[:, None]
    chs = np.arange(nch, dtype=np.int64)[None, :]
    out = np.empty((ncol, nch, 3))
    out[..., 0] = 1.5
    out[..., 1] = -0.2 + 1.1 * uniform(seed, _F_NDVI, cols, chs)
    out[..., 2] = 4.0 + 4.0 * uniform(seed, _F_BPDFC, cols, chs)
    return out


def make_ocean(ncol: int, channels_nm: Sequence[float], seed: int = 7, col0: int = 0):
    """Ocean surface inputs of the Cox-Munk drivers: 10 m wind components
    u10m, v10m [ncol] (m/s, |w| up to 14) and the water refractive index
    mr [nch] (1.34 at 400 nm falling slowly with wavelength)."""
    cols = np.arange(col0, col0 + ncol, dtype=np.int64)
    u = -10.0 + 20.0 * uniform(seed, _F_U10, cols)
    v = -10.0 + 20.0 * uniform(seed, _F_V10, cols)
    mr = 1.34 - 1.0e-5 * (np.asarray(channels_nm, dtype=np.float64) - 400.0)
    return u, v, mr


# Named workloads.  nstreams is the per-hemisphere stream count NSTREAMS.
WORKLOADS: Dict[str, dict] = {
    "scalar_1k_1ch_n4": dict(ncol=1024, km=72, channels=(550.0,), nstreams=4, nmom=32, npol=1, vector=False),
    "scalar_500k_4ch": dict(ncol=500_000, km=72, channels=(354.0, 388.0, 550.0, 870.0), nstreams=6, nmom=64,
                             npol=1, vector=False),
    "vector_200k_2ch": dict(ncol=200_000, km=72, channels=(550.0, 870.0), nstreams=4, nmom=32, npol=6,
                             vector=True),
    "uneven_300k_1ch": dict(ncol=300_000, km=72, channels=(550.0,), nstreams=None, nmom=32, npol=1, vector=False,
                             uneven_layers=(30, 72), stream_choices=(2, 4, 6, 8)),
}


# ---------------------------------------------------------------------------
# Device-side generator (torch) for large workloads.  The integer hash is the
# same SplitMix64 chain as ``uniform`` (two's-complement int64 arithmetic with
# logical shifts), so the uniforms are bit-identical to the host generator; the
# float transforms run on the GPU and agree with the host generator to within
# float rounding.

def _to_i64(v: int) -> int:
    v &= 0xFFFFFFFFFFFFFFFF
    return v - (1 << 64) if v >= (1 << 63) else v


def _srl(x, k: int):
    """Logical right shift of an int64 tensor."""
    import torch
    return (x >> k) & ((1 << (64 - k)) - 1)


def _splitmix64_t(x):
    x = x + _to_i64(0x9E3779B97F4A7C15)
    x = x ^ _srl(x, 30)
    x = x * _to_i64(0xBF58476D1CE4E5B9)
    x = x ^ _srl(x, 27)
    x = x * _to_i64(0x94D049BB133111EB)
    x = x ^ _srl(x, 31)
    return x


def uniform_t(seed: int, field: int, *idx):
    """torch version of ``uniform`` (idx are int64 tensors, broadcast)."""
    import torch
    key = np.uint64((seed * 0x100000001B3 + field * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    h = int(_splitmix64(np.asarray(key, dtype=np.uint64)))
    arrs = torch.broadcast_tensors(*idx)
    out = torch.full(arrs[0].shape, _to_i64(h), dtype=torch.int64, device=arrs[0].device)
    for a in arrs:
        out = _splitmix64_t(out ^ a)
    return _srl(out, 11).to(torch.float64) * (1.0 / 9007199254740992.0)


def make_columns_torch(ncol: int, km: int, channels_nm: Sequence[float], seed: int = 7, col0: int = 0,
                       device="cuda", uneven_layers: Tuple[int, int] | None = None) -> Dict[str, object]:
    """Device tensors of ``make_columns`` (same shapes, reversed-Fortran C order)."""
    import torch
    dev = torch.device(device)
    f32, f64 = torch.float32, torch.float64
    cols = torch.arange(col0, col0 + ncol, dtype=torch.int64, device=dev)
    c2 = cols[:, None]
    kk = torch.arange(km, device=dev, dtype=torch.int64)[None, :]
    nq = len(TRACERS)
    delp = (200.0 + 1800.0 * uniform_t(seed, _F_DELP, c2, kk)).to(f32)
    c3 = cols[:, None, None]
    q3 = torch.arange(nq, device=dev, dtype=torch.int64)[None, :, None]
    k3 = torch.arange(km, device=dev, dtype=torch.int64)[None, None, :]
    u1 = uniform_t(seed, _F_Q1, c3, q3, k3)
    u2 = uniform_t(seed, _F_Q2, c3, q3, k3)
    z = torch.sqrt(-2.0 * torch.log1p(-u1)) * torch.cos(2.0 * math.pi * u2)
    del u1, u2
    q = 1.0e-9 * torch.exp(math.log(1.5) * z)
    del z
    qm = (q.to(f32) * delp[:, None, :] / torch.tensor(9.81, dtype=f32, device=dev)).to(f32)
    del q
    rh = (0.99 * uniform_t(seed, _F_RH, c2, kk)).to(f32)
    # edge variables (getEdgeVars, float32)
    pe = torch.empty((ncol, km + 1), dtype=f32, device=dev)
    pe[:, 0] = 1.0
    for k in range(km):
        pe[:, k + 1] = pe[:, k] + delp[:, k]
    pm_true = 0.5 * (pe[:, :-1].to(f64) + pe[:, 1:].to(f64))
    tstd = torch.clamp(288.15 * (pm_true / 101325.0) ** 0.190263, min=216.65)
    airdens = (pm_true / (287.0 * tstd)).to(f32)
    grav = torch.tensor(9.80616, dtype=f32, device=dev)
    ze = torch.empty((ncol, km + 1), dtype=f32, device=dev)
    ze[:, km] = 0.0
    for k in range(km - 1, -1, -1):
        ze[:, k] = ze[:, k + 1] + delp[:, k] / (airdens[:, k] * grav)
    pmid = (pe[:, :-1] + pe[:, 1:]) / 2.0
    tm = pmid / (airdens * torch.tensor(287.0, dtype=f32, device=dev))
    te = torch.empty((ncol, km + 1), dtype=f32, device=dev)
    te[:, 0] = tm[:, 0]
    for k in range(1, km):
        a = torch.log(pe[:, k] / pmid[:, k - 1]) / torch.log(pmid[:, k] / pmid[:, k - 1])
        te[:, k] = tm[:, k - 1] + a * (tm[:, k] - tm[:, k - 1])
    te[:, km] = tm[:, km - 1] * (pe[:, km] / pmid[:, km - 1]) ** (2.0 / 7.0)
    sza = 70.0 * uniform_t(seed, _F_SZA, cols)
    vza = 60.0 * uniform_t(seed, _F_VZA, cols)
    raa = 180.0 * uniform_t(seed, _F_RAA, cols)
    ch = [float(w) for w in channels_nm]
    dp64 = delp.to(f64)
    frac = dp64 / dp64.sum(dim=1, keepdim=True)
    a550 = _rayleigh_xsec(0.55)
    rot = torch.stack([0.097 * frac * (_rayleigh_xsec(w * 1e-3) / a550) for w in ch], dim=0).contiguous()
    nch = len(ch)
    if uneven_layers is None:
        nl = torch.full((ncol,), km, dtype=torch.int32, device=dev)
    else:
        lo, hi = uneven_layers
        nl = (lo + torch.floor((hi - lo + 1) * uniform_t(seed, _F_NLAY, cols))).to(torch.int32)
    return dict(km=km, channels=np.asarray(ch), qm=qm.contiguous(), rh=rh.contiguous(), rot=rot,
                depol=torch.full((nch,), 0.03, dtype=f64, device=dev), pe=pe.to(f64), he=ze.to(f64), te=te.to(f64),
                albedo=torch.full((nch, ncol), 0.05, dtype=f64, device=dev), sza=sza, vza=vza, raa=raa,
                flux_factor=torch.ones((ncol, nch), dtype=f64, device=dev), nlayers=nl)
