"""Granule serving loop with continuous batching on one GPU.

Swath granules (a number of columns, 72 layers, 1-4 channels) arrive over
time.  The optics tables stay resident on the device; an admission controller
starts a granule as soon as a slot (a worker with its own HIP stream and work
arena) is free and the HBM budget has room for the granule's buffers; granules
running on different slots overlap on the GPU (their kernels interleave), and
a granule that arrives while others are running is admitted without waiting
for them to drain.  Each finished granule records its latency (completion -
arrival).

Column data: a device-resident pool of seeded columns (generated once) is the
"arriving" data; admitting a granule copies its columns out of the pool into
the granule's own input buffers on the granule's stream (the device-side
stand-in for the host->device upload of a real granule).
"""
from __future__ import annotations

import dataclasses
import math
import queue
import threading
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from . import synthetic as S


@dataclasses.dataclass
class Granule:
    gid: int
    ncol: int
    channels: Sequence[float]
    col0: int            # first column of the pool the granule's data comes from
    arrival: float = 0.0
    start: float = 0.0
    finish: float = 0.0
    radiance: Optional[np.ndarray] = None
    nfourier: Optional[np.ndarray] = None
    error: Optional[str] = None   # set when the granule failed; its slot keeps serving


class ServingError(RuntimeError):
    """One or more granules failed; ``failed`` lists them."""

    def __init__(self, failed: List[Granule]):
        super().__init__(f"{len(failed)} granule(s) failed: " +
                         "; ".join(f"granule {g.gid}: {g.error}" for g in failed[:5]))
        self.failed = failed


def granule_bytes(ncol: int, km: int, nch: int) -> int:
    """Device bytes of one granule's own buffers (inputs + outputs)."""
    nq = len(S.TRACERS)
    inp = ncol * (4 * km * nq + 4 * km + 8 * km * nch + 3 * 8 * (km + 1) + 3 * 8 + 8 * nch + 8 * nch)
    out = ncol * nch * (8 + 8 + 4)
    return inp + out


class ColumnPool:
    """Device-resident seeded columns (all four workload channels)."""

    def __init__(self, torch, ncol: int, km: int, channels: Sequence[float], device, seed: int = 7):
        self.torch = torch
        self.channels = list(channels)
        self.km = km
        self.ncol = ncol
        parts = []
        step = 100_000
        for c0 in range(0, ncol, step):
            parts.append(S.make_columns_torch(min(step, ncol - c0), km, channels, seed=seed, col0=c0, device=device))
        cat = lambda k, dim: torch.cat([p[k] for p in parts], dim=dim).contiguous()
        self.qm, self.rh = cat("qm", 0), cat("rh", 0)
        self.rot, self.albedo = cat("rot", 1), cat("albedo", 1)
        self.pe, self.he, self.te = cat("pe", 0), cat("he", 0), cat("te", 0)
        self.sza, self.vza, self.raa = cat("sza", 0), cat("vza", 0), cat("raa", 0)
        self.depol = parts[0]["depol"]


class Server:
    """Admission + continuous batching over `slots` concurrent GPU streams."""

    def __init__(self, torch, mie, pool: ColumnPool, nstreams: int, nmom: int, slots: int = 4,
                 hbm_budget: float = 200e9, arena_bytes: float = 28e9, keep_results: bool = False,
                 precision: int = 0):
        self.torch = torch
        self.precision = precision
        self.mie = mie
        self.pool = pool
        self.nstreams = nstreams
        self.nmom = nmom
        self.keep = keep_results
        self.dev = torch.device("cuda:0")
        # every slot holds a work + optics arena; the rest of the budget is for granule buffers
        self.slots = max(1, min(slots, int(hbm_budget // (arena_bytes + 1))))
        self.budget = hbm_budget - self.slots * arena_bytes
        self.in_use = 0.0
        self.cv = threading.Condition()
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(self.slots)]   # default priority
        self.tr_idx = torch.tensor([mie.query_idx(n) for n in S.TRACERS], dtype=torch.int32, device=self.dev)
        self.done: List[Granule] = []
        self.free_slots = list(range(self.slots))

    # -- one granule on one slot
    def _run(self, g: Granule, slot: int):
        torch = self.torch
        st = self.streams[slot]
        P = self.pool
        km = P.km
        chans = [P.channels.index(c) for c in g.channels]
        nch = len(chans)
        sl = slice(g.col0, g.col0 + g.ncol)
        with torch.cuda.stream(st):
            # the granule's own buffers (arrival of its data)
            qm = P.qm[sl].clone()
            rh = P.rh[sl].clone()
            rot = P.rot[chans][:, sl].contiguous()
            alb = P.albedo[chans][:, sl].contiguous()
            pe, he, te = P.pe[sl].clone(), P.he[sl].clone(), P.te[sl].clone()
            sza, vza, raa = P.sza[sl].clone(), P.vza[sl].clone(), P.raa[sl].clone()
            depol = P.depol[chans].contiguous()
            ch_idx = torch.tensor(chans, dtype=torch.int32, device=self.dev)
            rad = torch.empty((nch, g.ncol), dtype=torch.float64, device=self.dev)
            refl = torch.empty_like(rad)
            nf = torch.empty((nch, g.ncol), dtype=torch.int32, device=self.dev)
            p = lambda t: t.data_ptr()
            e = _lib.lib().aa_column_radiance_dev(
                self.mie.handle, km, g.ncol, nch, len(S.TRACERS), p(self.tr_idx), p(ch_idx), p(qm), p(rh),
                self.nmom, 1, self.nstreams, 0, self.precision, None, None, p(rot), p(depol), p(pe), p(he), p(te), p(alb),
                p(sza), p(raa), p(vza), None, S.MISSING, p(rad), p(refl), None, None, p(nf), st.cuda_stream)
            _lib.check(e, "aa_column_radiance_dev")
            if self.keep:
                g.radiance = rad.cpu().numpy()
                g.nfourier = nf.cpu().numpy()
            st.synchronize()
        g.finish = time.perf_counter()

    def _worker(self, slot: int, work: "queue.Queue[Optional[Granule]]"):
        while True:
            item = work.get()
            if item is None:
                return
            g, nbytes = item
            try:
                self._run(g, slot)
            except Exception as e:  # recorded on the granule; the slot keeps serving
                g.error = f"{type(e).__name__}: {e}"
                g.finish = time.perf_counter()
            finally:
                with self.cv:
                    self.in_use -= nbytes
                    self.free_slots.append(slot)
                    self.done.append(g)
                    self.cv.notify_all()

    def serve(self, granules: List[Granule], arrivals: Sequence[float], raise_on_error: bool = True) -> Dict[str, float]:
        """Serves granules arriving at the given offsets (seconds from start).

        A granule whose run raises gets ``error`` set and does not stop the
        loop; at the end serve() raises ServingError listing the failed
        granules (or, with raise_on_error=False, reports their count)."""
        with self.cv:
            self.done = []
        queues = [queue.Queue() for _ in range(self.slots)]
        threads = [threading.Thread(target=self._worker, args=(k, queues[k]), daemon=True) for k in range(self.slots)]
        for t in threads:
            t.start()
        t0 = time.perf_counter()
        pending: List[Granule] = []
        i = 0
        n = len(granules)
        started = 0
        while started < n:
            now = time.perf_counter() - t0
            while i < n and arrivals[i] <= now:
                granules[i].arrival = t0 + arrivals[i]
                pending.append(granules[i])
                i += 1
            with self.cv:
                # admit in arrival order while a slot and HBM budget are free
                while pending and self.free_slots:
                    g = pending[0]
                    nb = granule_bytes(g.ncol, self.pool.km, len(g.channels))
                    if self.in_use + nb > self.budget and self.in_use > 0:
                        break
                    pending.pop(0)
                    self.in_use += nb
                    slot = self.free_slots.pop(0)
                    g.start = time.perf_counter()
                    queues[slot].put((g, nb))
                    started += 1
                if started < n:
                    wait = (t0 + arrivals[i] - time.perf_counter()) if (i < n and not pending) else 0.05
                    self.cv.wait(timeout=max(0.001, min(wait, 0.05)))
        with self.cv:
            while len(self.done) < n:
                self.cv.wait(timeout=0.1)
        for q in queues:
            q.put(None)
        for t in threads:
            t.join()
        t1 = time.perf_counter()
        failed = [g for g in granules if g.error is not None]
        if failed and raise_on_error:
            raise ServingError(failed)
        lat = np.array([g.finish - g.arrival for g in granules])
        cols = sum(g.ncol for g in granules)
        span = max(g.finish for g in granules) - min(g.arrival for g in granules)
        return {"granules": n, "columns": int(cols), "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "throughput_cols_per_s": cols / span, "wall_s": t1 - t0, "slots": self.slots,
                "failed": len(failed)}


def make_granules(n: int, pool_ncol: int, channels: Sequence[float], seed: int = 11, lo: int = 1000,
                  hi: int = 200_000) -> List[Granule]:
    """Granules with column counts log-uniform in [lo, hi] and 1-4 channels,
    each a slice of the column pool."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        ncol = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        ncol = min(ncol, pool_ncol)
        nch = int(rng.integers(1, len(channels) + 1))
        chans = sorted(rng.choice(len(channels), size=nch, replace=False).tolist())
        col0 = int(rng.integers(0, pool_ncol - ncol + 1))
        out.append(Granule(gid=k, ncol=ncol, channels=[channels[c] for c in chans], col0=col0))
    return out


def poisson_arrivals(n: int, rate: float, seed: int = 12) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n)) - 0.0
