"""GPU-side helpers of the test-suite: the fused device path on torch buffers."""
import numpy as np

from aeroapps_amd import _lib
from aeroapps_amd import synthetic as S


def fused(torch, mie, cols, channels, N, nmom, npol, nlayers=None, nstreams_col=None, vector=False, stream=None,
          precision=0):
    """Runs aa_column_radiance_dev on host Columns; returns (rad, refl, nf[, q, u]) as
    reversed-Fortran [nch, nobs] arrays."""
    dev = torch.device("cuda:0")
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    f32, f64 = torch.float32, torch.float64
    nch, ncol, km = len(channels), cols.nobs, cols.km
    qm, rh = T(cols.qm, f32), T(cols.rh, f32)
    rot, depol = T(cols.rot, f64), T(cols.depol, f64)
    pe, he, te = T(cols.pe, f64), T(cols.he, f64), T(cols.te, f64)
    alb, sza, raa, vza = T(cols.albedo, f64), T(cols.sza, f64), T(cols.raa, f64), T(cols.vza, f64)
    ff = T(cols.flux_factor, f64)
    tr = T(np.array([mie.query_idx(n) for n in S.TRACERS]), torch.int32)
    ci = T(np.arange(nch), torch.int32)
    rad = torch.zeros((nch, ncol), dtype=f64, device=dev)
    refl = torch.zeros_like(rad)
    q = torch.zeros_like(rad) if vector else None
    u = torch.zeros_like(rad) if vector else None
    nf = torch.zeros((nch, ncol), dtype=torch.int32, device=dev)
    nl = T(nlayers, torch.int32) if nlayers is not None else None
    ns = T(nstreams_col, torch.int32) if nstreams_col is not None else None
    p = lambda t: t.data_ptr() if t is not None else None
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    with torch.cuda.stream(st):
        e = _lib.lib().aa_column_radiance_dev(mie.handle, km, ncol, nch, len(S.TRACERS), p(tr), p(ci), p(qm), p(rh),
                                              nmom, npol, N, int(vector), int(precision), p(nl), p(ns), p(rot), p(depol), p(pe),
                                              p(he), p(te), p(alb), p(sza), p(raa), p(vza), p(ff), S.MISSING, p(rad),
                                              p(refl), p(q), p(u), p(nf), st.cuda_stream)
    _lib.check(e, "aa_column_radiance_dev")
    torch.cuda.synchronize()
    out = [rad.cpu().numpy(), refl.cpu().numpy(), nf.cpu().numpy()]
    if vector:
        out += [q.cpu().numpy(), u.cpu().numpy()]
    return tuple(out)
