"""Python binding of the device-resident Chem_Mie optics tables and the batch
optics query ``VLIDORT_getAOPvector`` / ``VLIDORT_getAOPscalar``
(src/Components/rtms/geo_vlidort/VLIDORT_MieMod.F90:10-336; f2py contract of
``mieobs.getAOPvector``, src/Shared/GMAO_Shared/Chem_Base/MieObs_py.F90:218-456).

Arrays are numpy arrays in their Fortran shapes (qm[km, nq, nobs], ...), float32.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from ._lib import check, lib, _fp, _ip


def _f32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _p(a, kind=_fp):
    return a.ctypes.data_as(kind)


class MieTables:
    """Chem_Mie object (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:65-99):
    species tables already interpolated to the channels, uploaded once and kept
    on the GPU until ``close``."""

    def __init__(self, channels_m: Sequence[float]):
        self.channels_m = np.asarray(channels_m, dtype=np.float32)
        h = ctypes.c_void_p()
        check(lib().aa_mie_create(ctypes.byref(h), len(self.channels_m), _p(self.channels_m)), "aa_mie_create")
        self.handle = h

    def add_table(self, rh, bext, bsca, bbck, g, refr, refi, pmom, bin_names: Sequence[str]):
        """One species table; arrays in Fortran shape (nrh, nch, nbin) and
        pmom (nrh, nch, nbin, nmom, npol)."""
        rh = _f32(rh)
        bext, bsca, bbck, g, refr, refi, pmom = map(_f32, (bext, bsca, bbck, g, refr, refi, pmom))
        nrh, _, nbin, nmom, npol = pmom.shape
        names = b"".join(n.encode().ljust(16)[:16] for n in bin_names)
        assert len(bin_names) == nbin
        check(lib().aa_mie_add_table(self.handle, nrh, nbin, nmom, npol, _p(rh), _p(bext), _p(bsca), _p(bbck),
                                     _p(g), _p(refr), _p(refi), _p(pmom), names), "aa_mie_add_table")

    def add_table_native(self, lambda_table, rh, reff, bext, bsca, bbck, g, bin_names: Sequence[str], nmom: int = 0,
                         pback=None, pmom=None, gf=None, rhop=None, refr=None, refi=None) -> int:
        """One species table from the variables of its optics-table file
        (Chem_MieTableRead, src/Shared/GMAO_Shared/Chem_Base/Chem_MieTableMod.F90:
        238-719): float64 arrays in the file's Fortran shapes, lambda_table[nch_table]
        [m], rh[nrh], reff/gf/rhop[nrh, nbin], bext..refi[nch_table, nrh, nbin],
        pback[nch_table, nrh, nbin, npol], pmom[nch_table, nrh, nbin, nmom_table,
        npol]; absent variables as None.  Keeps ``nmom`` moments.  Returns rc (99:
        the file has fewer moments than asked)."""
        from ._lib import _dp
        f64 = lambda a: None if a is None else np.asfortranarray(np.asarray(a, dtype=np.float64))
        lam, rh, reff, bext, bsca, bbck, g, pback, pmom, gf, rhop, refr, refi = map(
            f64, (lambda_table, rh, reff, bext, bsca, bbck, g, pback, pmom, gf, rhop, refr, refi))
        nch_t, nrh, nbin = bext.shape
        nmom_t = pmom.shape[3] if pmom is not None else 0
        npol = pmom.shape[4] if pmom is not None else (pback.shape[3] if pback is not None else 6)
        names = b"".join(n.encode().ljust(16)[:16] for n in bin_names)
        assert len(bin_names) == nbin
        pd = lambda a: None if a is None else a.ctypes.data_as(_dp)
        rc = np.zeros(1, dtype=np.int32)
        check(lib().aa_mie_add_table_native(self.handle, nch_t, pd(lam), nrh, nbin, nmom_t, npol, int(nmom), pd(rh),
                                            pd(reff), pd(bext), pd(bsca), pd(bbck), pd(g), pd(pback), pd(pmom),
                                            pd(gf), pd(rhop), pd(refr), pd(refi), names, _p(rc, _ip)),
              "aa_mie_add_table_native")
        return int(rc[0])

    def query_extra(self, idx: int, channel: float, rh):
        """rEff [um], p11, p22, gf, rhod, vol, area of Chem_MieQuery for arrays of
        relative humidity (tables added with ``add_table_native``)."""
        r = np.ascontiguousarray(rh, dtype=np.float32).ravel()
        keys = ("reff", "p11", "p22", "gf", "rhod", "vol", "area")
        outs = {k: np.zeros(len(r), dtype=np.float32) for k in keys}
        check(lib().aa_mie_query_extra(self.handle, idx, float(channel), len(r), _p(r), *[_p(outs[k]) for k in keys]),
              "aa_mie_query_extra")
        return outs

    @classmethod
    def from_rc(cls, rcfile: str, reader, chem_reg=None):
        """Chem_MieCreate(rcfile, rc[, chemReg]) (src/Shared/GMAO_Shared/Chem_Base/
        Chem_MieMod.F90:124-411) in the library: the resource file and the
        registry are parsed there, the seven optics-table files are read by
        ``reader(path)``, a callable returning the file's variables as a dict
        of float64 arrays in Fortran shapes (lambda_table, rh, reff, bext, bsca,
        bbck, g and the optional pback, pmom, gf, rhop, refr, refi, as for
        ``add_table_native``).  ``chem_reg`` is (names, species) as from
        ``registry()``; by default Chem_MieRegistry.rc, else Chem_Registry.rc,
        in the working directory.  Returns (MieTables or None, rc)."""
        import os
        from ._lib import _vp
        if chem_reg is None:
            path = "Chem_MieRegistry.rc" if os.path.exists("Chem_MieRegistry.rc") else "Chem_Registry.rc"
            names, species, rc = registry(path)
            if rc:
                return None, 3
        else:
            names, species = chem_reg
        nq = len(names)
        cnames = b"".join(n.encode().ljust(16)[:16] for n in names) or b" " * 16
        csp = np.asarray(list(species) or [-1], dtype=np.int32)
        keep = []

        class TableFile(ctypes.Structure):
            _fields_ = [(k, ctypes.c_int) for k in ("nch", "nrh", "nbin", "nmom", "npol")] + \
                       [(k, ctypes.c_void_p) for k in ("lambda_", "rh", "reff", "bext", "bsca", "bbck", "g", "pback",
                                                       "pmom", "gf", "rhop", "refr", "refi")]

        proto = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(TableFile))

        def cb(user, path, tf):
            try:
                d = reader(path.decode())
            except Exception:
                return 1
            f64 = lambda a: None if a is None else np.asfortranarray(np.asarray(a, dtype=np.float64))
            arrs = {k: f64(d.get(k)) for k in ("lambda_table", "rh", "reff", "bext", "bsca", "bbck", "g", "pback",
                                               "pmom", "gf", "rhop", "refr", "refi")}
            keep.append(arrs)
            t = tf.contents
            t.nch, t.nrh, t.nbin = arrs["bext"].shape
            t.nmom = arrs["pmom"].shape[3] if arrs["pmom"] is not None else 0
            t.npol = arrs["pmom"].shape[4] if arrs["pmom"] is not None else (
                arrs["pback"].shape[3] if arrs["pback"] is not None else 0)
            ptr = lambda a: None if a is None else a.ctypes.data
            t.lambda_ = ptr(arrs["lambda_table"])
            for k in ("rh", "reff", "bext", "bsca", "bbck", "g", "pback", "pmom", "gf", "rhop", "refr", "refi"):
                setattr(t, k, ptr(arrs[k]))
            return 0

        cfun = proto(cb)
        h = ctypes.c_void_p()
        rc = np.zeros(1, dtype=np.int32)
        check(lib().aa_mie_create_rc(ctypes.byref(h), rcfile.encode(), nq, cnames, _p(csp, _ip),
                                     ctypes.cast(cfun, _vp), None, _p(rc, _ip)), "aa_mie_create_rc")
        if rc[0] != 0:
            return None, int(rc[0])
        t = cls.__new__(cls)
        t.handle = h
        nch = t.dims()[0]
        t.channels_m = np.zeros(nch, dtype=np.float32)
        check(lib().aa_mie_channels(h, _p(t.channels_m)), "aa_mie_channels")
        return t, 0

    def tracer_table(self, idx: int):
        """(table, bin, nbin, nrh, nmom, npol) of tracer idx (Chem_MieQueryIdx's
        vindex / vtableUse)."""
        v = [ctypes.c_int() for _ in range(6)]
        check(lib().aa_mie_tracer_table(self.handle, idx, *[ctypes.byref(x) for x in v]), "aa_mie_tracer_table")
        return tuple(x.value for x in v)

    @classmethod
    def from_synthetic(cls, tables) -> "MieTables":
        """Tables built by aeroapps_amd.synthetic.make_mie_tables."""
        t = cls(tables[0].lambda_m)
        for tb in tables:
            t.add_table(tb.rh, tb.bext.T, tb.bsca.T, tb.bbck.T, tb.g.T, tb.refr.T, tb.refi.T, tb.pmom.T, tb.tracers)
        return t

    def query_idx(self, qname: str) -> int:
        return lib().aa_mie_query_idx(self.handle, qname.encode())

    def dims(self):
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().aa_mie_dims(self.handle, *[ctypes.byref(x) for x in v]), "aa_mie_dims")
        return tuple(x.value for x in v)

    def query(self, idx: int, channel: float, q_mass, rh, nmom: int = 0, npol: int = 1):
        """Chem_MieQuery over arrays of (q_mass, rh) for one tracer index and
        1-based channel; returns dict of tau, ssa, gasym, bext, bsca, bbck,
        refr, refi and pmom (n, nmom, npol) when nmom > 0."""
        q = np.ascontiguousarray(q_mass, dtype=np.float32).ravel()
        r = np.ascontiguousarray(rh, dtype=np.float32).ravel()
        n = len(q)
        outs = {k: np.zeros(n, dtype=np.float32) for k in ("tau", "ssa", "gasym", "bext", "bsca", "bbck", "refr",
                                                             "refi")}
        pm = np.zeros((n, nmom, npol), dtype=np.float32, order="F") if nmom > 0 else None
        check(lib().aa_mie_query(self.handle, idx, float(channel), n, _p(q), _p(r),
                                 *[_p(outs[k]) for k in ("tau", "ssa", "gasym", "bext", "bsca", "bbck", "refr",
                                                         "refi")], nmom, npol,
                                 _p(pm) if pm is not None else None), "aa_mie_query")
        if pm is not None:
            outs["pmom"] = pm
        return outs

    def close(self):
        if self.handle:
            lib().aa_mie_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _vname(vnames: Sequence[str]) -> bytes:
    nq = len(vnames)
    buf = bytearray(16 * nq)
    for iq, nm in enumerate(vnames):
        s = nm.encode().ljust(16)[:16]
        for c in range(16):
            buf[c * nq + iq] = s[c]   # Fortran character array vname(nq,16)
    return bytes(buf)


def get_aop_vector(mie: MieTables, channels, vname: Sequence[str], verbose, qm, rh, nMom: int, nPol: int):
    """VLIDORT_getAOPvector: qm(km,nq,nobs), rh(km,nobs) ->
    tau, ssa, g (km,nch,nobs), pmom (km,nch,nobs,nMom,nPol), rc."""
    ch = _f32(channels)
    qm, rh = _f32(qm), _f32(rh)
    km, nq, nobs = qm.shape
    nch = len(ch)
    tau = np.zeros((km, nch, nobs), dtype=np.float32, order="F")
    ssa = np.zeros_like(tau)
    g = np.zeros_like(tau)
    pm = np.zeros((km, nch, nobs, nMom, nPol), dtype=np.float32, order="F")
    rc = np.zeros(1, dtype=np.int32)
    check(lib().aa_get_aop_vector(mie.handle, km, nobs, nch, nq, _p(ch), _vname(vname), int(verbose), _p(qm), _p(rh),
                                  nMom, nPol, _p(tau), _p(ssa), _p(g), _p(pm), _p(rc, _ip)), "aa_get_aop_vector")
    return tau, ssa, g, pm, int(rc[0])


def get_aop_scalar(mie: MieTables, channels, vname: Sequence[str], verbose, qm, rh):
    """VLIDORT_getAOPscalar: -> tau, ssa, g (km,nch,nobs), rc."""
    ch = _f32(channels)
    qm, rh = _f32(qm), _f32(rh)
    km, nq, nobs = qm.shape
    nch = len(ch)
    tau = np.zeros((km, nch, nobs), dtype=np.float32, order="F")
    ssa = np.zeros_like(tau)
    g = np.zeros_like(tau)
    rc = np.zeros(1, dtype=np.int32)
    check(lib().aa_get_aop_scalar(mie.handle, km, nobs, nch, nq, _p(ch), _vname(vname), int(verbose), _p(qm), _p(rh),
                                  _p(tau), _p(ssa), _p(g), _p(rc, _ip)), "aa_get_aop_scalar")
    return tau, ssa, g, int(rc[0])


def getEdgeVars(aer, ptop=1.0):
    """mieobs.getEdgeVars (src/Components/misc/scat/mieobs.py:60-76 ->
    getEdgeVars, src/Shared/GMAO_Shared/Chem_Base/MieObs_py.F90:27-83): edge
    pressure [Pa], height above the surface [m] and temperature [K], each
    [km+1, nobs], from ``aer.AIRDENS`` and ``aer.DELP`` given as (km, nobs) or
    (nobs, km) float32 arrays."""
    ad, dp = np.asarray(aer.AIRDENS, dtype=np.float32), np.asarray(aer.DELP, dtype=np.float32)
    if needs_transpose(aer):
        ad, dp = ad.T, dp.T
    return edge_vars(ad, dp, ptop)


def needs_transpose(aer) -> bool:
    """True when the arrays of ``aer`` are (nobs, km): the surface pressure
    ``aer.PS`` has as many entries as DELP's first dimension
    (src/Components/misc/scat/mieobs.py:353-364)."""
    return np.shape(aer.PS)[0] == np.shape(aer.DELP)[0]


def edge_vars(airdens, delp, ptop=1.0):
    """getEdgeVars on (km, nobs) float32 arrays; returns (pe, ze, te)."""
    ad, dp = _f32(airdens), _f32(delp)
    km, nobs = ad.shape
    pe = np.zeros((km + 1, nobs), dtype=np.float32, order="F")
    ze = np.zeros_like(pe)
    te = np.zeros_like(pe)
    check(lib().aa_edge_vars(km, nobs, _p(ad), _p(dp), float(ptop), _p(pe), _p(ze), _p(te)), "aa_edge_vars")
    return pe, ze, te


def calc_qm(q, delp):
    """calc_qm (src/Components/rtms/geo_vlidort/geo_vlidort.F90:2124-2144):
    qm[km, nq, nobs] = q * delp / 9.81 (float32), q[km, nq, nobs], delp[km, nobs]."""
    q, dp = _f32(q), _f32(delp)
    km, nq, nobs = q.shape
    qm = np.zeros((km, nq, nobs), dtype=np.float32, order="F")
    check(lib().aa_calc_qm(km, nq, nobs, _p(q), _p(dp), _p(qm)), "aa_calc_qm")
    return qm


def parse_mie_rc(text: str) -> dict:
    """The resource file Chem_MieCreateFromRC reads with i90
    (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:200-300): ``label: value``
    lines, ``#`` comments, tables of values after ``label::`` until ``::``.
    Returns the labels with their value strings (tables as lists of rows)."""
    out, table, rows = {}, None, []
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if table is not None:
            if line == "::":
                out[table] = rows
                table, rows = None, []
            elif line:
                rows.append(line.split())
            continue
        if not line:
            continue
        if line.endswith("::"):
            table = line[:-2].strip()
            continue
        if ":" in line:
            k, v = line.split(":", 1)
            out[k.strip()] = v.strip()
    return out


def registry(path: str):
    """Chem_RegistryCreate (src/Shared/GMAO_Shared/Chem_Base/Chem_RegistryMod.F90:
    200-545) in the library: (tracer names, species codes AA_SP_* or -1, rc) of
    the registry file ``path``."""
    nq = ctypes.c_int()
    rc = ctypes.c_int()
    check(lib().aa_mie_registry(path.encode(), 0, ctypes.byref(nq), None, None, ctypes.byref(rc)), "aa_mie_registry")
    if rc.value:
        return [], [], rc.value
    n = nq.value
    buf = ctypes.create_string_buffer(16 * max(n, 1))
    sp = np.zeros(max(n, 1), dtype=np.int32)
    check(lib().aa_mie_registry(path.encode(), n, ctypes.byref(nq), buf, _p(sp, _ip), ctypes.byref(rc)),
          "aa_mie_registry")
    names = [buf.raw[16 * i:16 * i + 16].decode().rstrip() for i in range(n)]
    return names, [int(x) for x in sp[:n]], rc.value


SPECIES_CODES = {"DU": 0, "SS": 1, "SU": 2, "OC": 3, "BC": 4, "NI": 5, "BRC": 6}


def mie_rc_settings(text: str) -> dict:
    """The settings Chem_MieCreateFromRC takes from a Chem_Mie resource file
    (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:200-300): ``n_channels``,
    ``n_moments`` (0 when absent), ``r_channels`` [m] and the optics-table file
    ``filename_optical_properties_XX`` of DU, SS, BC, OC, SU, NI, BRC."""
    d = parse_mie_rc(text)
    nch = int(d["n_channels"])
    ch = [float(v) for v in d["r_channels"].replace(",", " ").split()][:nch]
    files = {sp: d.get(f"filename_optical_properties_{sp}") for sp in ("DU", "SS", "BC", "OC", "SU", "NI", "BRC")}
    return dict(n_channels=nch, channels_m=ch, nmom=int(d.get("n_moments", 0)), optics_files=files)


def species_bin_names(species: str, tracer_names: Sequence[str]) -> list:
    """Tracers a species table serves, bin by bin, as Chem_MieCreateFromRC
    assigns them (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:338-400):
    every tracer of the species in order, except that the sulfate table serves
    only names starting so4/SO4/sul/SUL and the nitrate table only no3/NO3."""
    names = [str(n) for n in tracer_names]
    if species.upper() == "SU":
        return [n for n in names if n[:3] in ("so4", "SO4", "sul", "SUL")]
    if species.upper() == "NI":
        return [n for n in names if n[:3] in ("no3", "NO3")]
    return names
