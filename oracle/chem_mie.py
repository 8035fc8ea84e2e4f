"""ORACLE TEST INFRASTRUCTURE -- not part of the product.

numpy float32 restatement of the Chem_Mie optics query and the per-tracer
optics summation.  Only tests/, __graft_entry__.smoke() and the bench's
cpu_baseline leg may use it, and only as the checker.

Every operation is an IEEE float32 operation in the reference's order
(numpy evaluates ``a*b + c*d`` as two rounded products and a rounded sum, as the
reference's flang/x86-64 build does; there is no fused multiply-add).

Reference (src/Shared/GMAO_Shared/Chem_Base = CB, src/Components/rtms/geo_vlidort = GEO):
  rh_map            CB/Chem_MieTableMod.F90:612-627   (991-entry RH hash rhi/rha)
  polint            CB/Chem_MieTableMod.F90:671-717   (linear wavelength interpolation)
  table_read        CB/Chem_MieTableMod.F90:238-719   (Chem_MieTableRead after the netCDF reads)
  query_extra       CB/Chem_MieMod.F90:885-943        (rEff, p11, p22, gf, rhod, vol, area)
  query             CB/Chem_MieMod.F90:789-965        (Chem_MieQueryByInt)
  query_idx         CB/Chem_MieMod.F90:720-777        (Chem_MieQueryIdx)
  get_aop_vector    GEO/VLIDORT_MieMod.F90:124-336    (VLIDORT_getAOPvector)
  get_aop_scalar    GEO/VLIDORT_MieMod.F90:10-120     (VLIDORT_getAOPscalar)
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

f32 = np.float32


def rh_map(rh: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """rhi(991) (1-based table indices) and rha(991) for an RH grid (float32).

    For slot j (0-based j = j_ref-1) the last i scanning from the top with
    j >= int(rh(i)*1000) is chosen; the slope is ((j)/1000. - rh(i)) /
    (rh(i+1) - rh(i)), 0 at the top node.  Slots with no such i keep 0
    (unset in the reference)."""
    rh = np.asarray(rh, dtype=f32)
    nrh = len(rh)
    rhi = np.zeros(991, dtype=np.int32)
    rha = np.zeros(991, dtype=f32)
    thresh = (rh * f32(1000.0)).astype(f32).astype(np.int64)   # int() truncates toward zero
    for j in range(991):
        for i in range(nrh - 1, -1, -1):
            if j >= thresh[i]:
                rhi[j] = i + 1
                if i + 1 >= nrh:
                    rha[j] = f32(0.0)
                else:
                    num = f32(f32(j) / f32(1000.0)) - rh[i]
                    rha[j] = f32(num) / f32(rh[i + 1] - rh[i])
                break
    return rhi, rha


def polint(x: np.ndarray, y: np.ndarray, xwant: float) -> np.float32:
    """Linear interpolation of float64 table data y(x) at float32 xwant, result
    float32 (CB/Chem_MieTableMod.F90:671-717): out-of-range clamps to the end
    node (slope 0 at the top node).  Fortran kinds: slope and dx are default
    reals assigned from float64 expressions (xWant is promoted in
    ``xWant - x(i)``), slope*dx is a float32 product, and ``y(i) + slope*dx``
    is a float64 sum rounded to float32 on assignment."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = len(x)
    xw = f32(xwant)
    xd = float(xw)
    i = 0
    if xd < x[0]:
        i = 1
    if xd > x[n - 1]:
        i = n
    if i == 0:
        for j in range(n):
            if xd >= x[j]:
                i = j + 1
    if i == n:
        slope = f32(0.0)
    else:
        slope = f32((y[i] - y[i - 1]) / (x[i] - x[i - 1]))
    dx = f32(xd - float(x[i - 1]))
    sd = f32(slope * dx)
    return f32(float(y[i - 1]) + float(sd))


def table_read(lambda_table, channels_m, rh, reff, bext, bsca, bbck, g, nmom=0, pback=None, pmom=None, gf=None,
               rhop=None, refr=None, refi=None):
    """What Chem_MieTableRead (CB/Chem_MieTableMod.F90:238-719) does after its
    netCDF reads.  Arrays in reversed-Fortran C order: rh[nrh], reff/gf/rhop
    [nbin, nrh], bext..refi[nbin, nrh, nch_table], pback[npol, nbin, nrh,
    nch_table], pmom[npol, nmom_table, nbin, nrh, nch_table]; absent ones None.
    Returns (rc, dict) with float32 tables bext..refi[nbin, nch, nrh],
    pback[npol, nbin, nch, nrh], pmom[npol, nmom, nbin, nch, nrh], rh,
    reff/gf/rhop/rhod/vol/area[nbin, nrh] and the RH map."""
    lam = np.asarray(lambda_table, dtype=np.float64)
    bext = np.asarray(bext, dtype=np.float64)
    nbin, nrh, nl = bext.shape
    if nmom > 0 and (pmom is None or nmom > np.asarray(pmom).shape[1]):
        return 99, None
    npol = np.asarray(pmom).shape[0] if pmom is not None else (np.asarray(pback).shape[0] if pback is not None
                                                                 else 6)
    full = lambda v: np.full((nbin, nrh, nl), v)
    pb = np.asarray(pback, dtype=np.float64) if pback is not None else np.ones((npol, nbin, nrh, nl))
    gfv = np.asarray(gf, dtype=np.float64) if gf is not None else np.full((nbin, nrh), -999.0)
    if rhop is not None:
        rhp = np.asarray(rhop, dtype=np.float64)
        rhd = np.repeat(rhp[:, :1], nrh, axis=1)          # rhod from 'rhop', first RH row
    else:
        rhp = np.full((nbin, nrh), -999.0)
        rhd = np.full((nbin, nrh), -999.0)
    rr = np.asarray(refr, dtype=np.float64) if refr is not None else full(-999.0)
    ri = np.abs(np.asarray(refi, dtype=np.float64)) if refi is not None else full(-999.0)
    reff = np.asarray(reff, dtype=np.float64)
    vol = gfv * gfv * gfv / rhd
    area = vol / (float(f32(4.0) / f32(3.0)) * reff)
    nch = len(channels_m)
    out = {k: np.zeros((nbin, nch, nrh), dtype=f32) for k in ("bext", "bsca", "bbck", "g", "refr", "refi")}
    out["pback"] = np.zeros((npol, nbin, nch, nrh), dtype=f32)
    out["pmom"] = np.zeros((npol, nmom, nbin, nch, nrh), dtype=f32)
    src = dict(bext=bext, bsca=np.asarray(bsca, np.float64), bbck=np.asarray(bbck, np.float64),
               g=np.asarray(g, np.float64), refr=rr, refi=ri)
    pm = np.asarray(pmom, dtype=np.float64) if nmom > 0 else None
    for j in range(nbin):
        for i in range(nrh):
            for c, cw in enumerate(channels_m):
                for k, v in src.items():
                    out[k][j, c, i] = polint(lam, v[j, i], cw)
                for p in range(npol):
                    out["pback"][p, j, c, i] = polint(lam, pb[p, j, i], cw)
                for m in range(nmom):
                    for p in range(npol):
                        out["pmom"][p, m, j, c, i] = polint(lam, pm[p, m, j, i], cw)
    out["rh"] = np.asarray(rh, dtype=np.float64).astype(f32)
    for k, v in (("reff", reff), ("gf", gfv), ("rhop", rhp), ("rhod", rhd), ("vol", vol), ("area", area)):
        out[k] = v.astype(f32)
    out["rhi"], out["rha"] = rh_map(out["rh"])
    return 0, out


def query_extra(tab, ibin: int, ich: int, rh):
    """rEff [um], p11, p22, gf, rhod, vol, area of Chem_MieQueryByInt
    (CB/Chem_MieMod.F90:885-943) from a table_read result."""
    nrh = len(tab["rh"])
    irh, irhp1, arh = snap(rh, tab["rhi"], tab["rha"], nrh)
    lin = lambda v: _lin(v[irh - 1], v[irhp1 - 1], arh)
    return dict(reff=(f32(1.0e6) * lin(tab["reff"][ibin])).astype(f32),
                p11=lin(tab["pback"][0, ibin, ich]), p22=lin(tab["pback"][4, ibin, ich]),
                gf=lin(tab["gf"][ibin]), rhod=np.full(len(irh), tab["rhod"][ibin, 0], dtype=f32),
                vol=lin(tab["vol"][ibin]), area=lin(tab["area"][ibin]))


def query_idx(table_names: Sequence[str], qname: str) -> int:
    """0-based tracer index of ``qname`` (case-insensitive, ``XX::`` prefix
    stripped) or -1."""
    name = qname.strip()
    if "::" in name:
        name = name[name.index("::") + 2:]
    up = name.strip().upper()
    for i, t in enumerate(table_names):
        if t.strip().upper() == up:
            return i
    return -1


def snap(rh: np.ndarray, rhi: np.ndarray, rha: np.ndarray, nrh: int):
    """(irh, irhp1, arh) for float32 RH values; irh/irhp1 are 1-based."""
    rh = np.asarray(rh, dtype=f32)
    rhuse = np.minimum(rh, f32(0.99))             # the reference discards max(rh,0.)
    isnap = ((rhuse + f32(0.001)).astype(f32) * f32(1000.0)).astype(f32)
    isnap = np.trunc(isnap).astype(np.int64)
    isnap = np.where(isnap < 1, 1, isnap)
    arh = rha[isnap - 1]
    irh = rhi[isnap - 1].astype(np.int64)
    irhp1 = np.minimum(irh + 1, nrh)
    return irh, irhp1, arh


def _lin(tab_lo, tab_hi, arh):
    one_m = (f32(1.0) - arh).astype(f32)
    return ((tab_lo * one_m).astype(f32) + (tab_hi * arh).astype(f32)).astype(f32)


def query(bext, bsca, g, pmom, rhi, rha, ibin: int, ich: int, q, rh):
    """Chem_MieQueryByInt for one tracer bin / channel over arrays of (q, rh).

    Table arrays are C-ordered reversed-Fortran: bext[nbin, nch, nrh],
    pmom[npol, nmom, nbin, nch, nrh].  Returns tau, ssa, g, pmom[..., nmom, npol]."""
    nrh = bext.shape[-1]
    irh, irhp1, arh = snap(rh, rhi, rha, nrh)
    b = bext[ibin, ich]
    s = bsca[ibin, ich]
    gt = g[ibin, ich]
    bext_in = _lin(b[irh - 1], b[irhp1 - 1], arh)
    bsca_in = _lin(s[irh - 1], s[irhp1 - 1], arh)
    g_in = _lin(gt[irh - 1], gt[irhp1 - 1], arh)
    pm = pmom[:, :, ibin, ich, :]                      # [npol, nmom, nrh]
    pm_lo = np.moveaxis(pm[:, :, irh - 1], (0, 1), (-1, -2))   # [..., nmom, npol]
    pm_hi = np.moveaxis(pm[:, :, irhp1 - 1], (0, 1), (-1, -2))
    a = arh[..., None, None]
    pm_in = ((pm_lo * (f32(1.0) - a).astype(f32)).astype(f32) + (pm_hi * a).astype(f32)).astype(f32)
    tau = (bext_in * np.asarray(q, dtype=f32)).astype(f32)
    ssa = (bsca_in / bext_in).astype(f32)
    return tau, ssa, g_in, pm_in


def get_aop_vector(tables, channels_nm: Sequence[float], mie_channels_m: np.ndarray, vnames: Sequence[str],
                   qm: np.ndarray, rh: np.ndarray, nmom: int, npol: int):
    """VLIDORT_getAOPvector.  qm[nobs, nq, km], rh[nobs, km] (float32).

    ``tables`` is a list of synthetic.MieTable.  Returns
    (tau, ssa, g, pmom, rc) with tau/ssa/g as [nobs, nch, km] and pmom as
    [npol, nmom, nobs, nch, km] (reversed Fortran shapes), float32."""
    nobs, nq, km = qm.shape
    chans = np.asarray(channels_nm, dtype=f32)
    nch = len(chans)
    # channel matching in float32 (GEO/VLIDORT_MieMod.F90:210-225)
    idxc = []
    for n in range(nch):
        found = -1
        for m, lam in enumerate(mie_channels_m):
            if abs(f32(chans[n] - f32(f32(1.0e9) * f32(lam)))) < f32(1.0):
                found = m
                break
        idxc.append(found)
    tau = np.zeros((nobs, nch, km), dtype=f32)
    ssa = np.zeros_like(tau)
    g = np.zeros_like(tau)
    pm = np.zeros((nobs, nch, km, nmom, npol), dtype=f32)
    if any(i < 0 for i in idxc):
        return tau, ssa, g, np.ascontiguousarray(np.moveaxis(pm, (3, 4), (1, 0))), 99
    # tracer -> (table, bin)
    lookup = {}
    for t in tables:
        for b, name in enumerate(t.tracers):
            lookup[name.upper()] = (t, b)
    names = [name for t in tables for name in t.tracers]
    rc = 0
    maps = {id(t): rh_map(t.rh) for t in tables}
    for iq in range(nq):
        k = query_idx(names, vnames[iq])
        rc = 1 if k < 0 else 0
        if k < 0:
            continue
        t, ibin = lookup[names[k].upper()]
        if nmom > t.nmom:
            return tau, ssa, g, np.ascontiguousarray(np.moveaxis(pm, (3, 4), (1, 0))), 666
        if npol <= t.npol:
            npol_, sph = npol, False
        elif npol == 6 and t.npol == 4:
            npol_, sph = 4, True
        else:
            return tau, ssa, g, np.ascontiguousarray(np.moveaxis(pm, (3, 4), (1, 0))), 777
        rhi, rha = maps[id(t)]
        for n in range(nch):
            tau_, ssa_, g_, pm_ = query(t.bext, t.bsca, t.g, t.pmom, rhi, rha, ibin, idxc[n],
                                        qm[:, iq, :], rh)
            tau[:, n] = (tau[:, n] + tau_).astype(f32)
            st = (ssa_ * tau_).astype(f32)
            ssa[:, n] = (ssa[:, n] + st).astype(f32)
            g[:, n] = (g[:, n] + ((g_ * ssa_).astype(f32) * tau_).astype(f32)).astype(f32)
            w = ((pm_[..., :nmom, :npol_] * ssa_[..., None, None]).astype(f32) * tau_[..., None, None]).astype(f32)
            pm[:, n, :, :, :npol_] = (pm[:, n, :, :, :npol_] + w).astype(f32)
            if sph:
                pm[:, n, :, :, 4] = pm[:, n, :, :, 0]
                pm[:, n, :, :, 5] = pm[:, n, :, :, 2]
    pos = tau > f32(0.0)
    ssa = np.where(pos, (ssa / np.where(pos, tau, f32(1.0))).astype(f32), ssa).astype(f32)
    st = (ssa * tau).astype(f32)
    posst = st > f32(0.0)
    g = np.where(posst, (g / np.where(posst, st, f32(1.0))).astype(f32), g).astype(f32)
    pm = np.where(posst[..., None, None], (pm / np.where(posst, st, f32(1.0))[..., None, None]).astype(f32),
                  pm).astype(f32)
    return tau, ssa, g, np.ascontiguousarray(np.moveaxis(pm, (3, 4), (1, 0))), rc


def get_aop_scalar(tables, channels_nm, mie_channels_m, vnames, qm, rh):
    """VLIDORT_getAOPscalar: tau, ssa, g without moments."""
    tau, ssa, g, _, rc = get_aop_vector(tables, channels_nm, mie_channels_m, vnames, qm, rh, 1, 1)
    return tau, ssa, g, rc
