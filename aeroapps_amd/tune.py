"""Launch-shape tuner of the radiance kernels.

For each (streams, Stokes, precision) class it times the fused device path
(``aa_column_radiance_dev``) on a seeded sample with every launch shape this
build has -- the band solver's lanes per item (``bvp_lanes``) and the SAB/DAB
kernel's (item, layer) pairs per wavefront (``sab_pairs``) -- keeps the
fastest, and writes the table the library reads at start-up
(``aeroapps_amd/tuning_gfx950.txt``; format in include/aeroapps_amd.h).  All
shapes give bit-identical results; the tuner checks that too.

    python -m aeroapps_amd.tune [--ncol 20000] [--out aeroapps_amd/tuning_gfx950.txt]
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from . import synthetic as S

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning_gfx950.txt")


class Tuning(ctypes.Structure):
    _fields_ = [("bvp_lanes", ctypes.c_int), ("sab_pairs", ctypes.c_int), ("chunk_items", ctypes.c_int)]


def set_tuning(nstreams: int, nstokes: int, precision: int, bvp_lanes: int = 0, sab_pairs: int = 0,
               chunk_items: int = 0) -> None:
    t = Tuning(bvp_lanes, sab_pairs, chunk_items)
    _lib.check(_lib.lib().aa_set_tuning(nstreams, nstokes, precision, ctypes.byref(t)), "aa_set_tuning")


def clear_tuning(nstreams: int, nstokes: int, precision: int) -> None:
    _lib.check(_lib.lib().aa_set_tuning(nstreams, nstokes, precision, None), "aa_set_tuning")


def get_tuning(nstreams: int, nstokes: int, precision: int) -> Dict[str, int]:
    t = Tuning()
    _lib.check(_lib.lib().aa_get_tuning(nstreams, nstokes, precision, ctypes.byref(t)), "aa_get_tuning")
    return {"bvp_lanes": t.bvp_lanes, "sab_pairs": t.sab_pairs, "chunk_items": t.chunk_items}


def load_tuning(path: Optional[str] = None) -> None:
    _lib.check(_lib.lib().aa_load_tuning(path.encode() if path else None), "aa_load_tuning")


def shape_candidates(nstreams: int, nstokes: int) -> List[Tuple[int, int]]:
    """(bvp_lanes, sab_pairs) pairs this build has for the class: the values
    the library accepts, as it reports them."""
    orig = get_tuning(nstreams, nstokes, 0)
    out = set()
    for g in (4, 8, 16, 32, 64):
        for p in (16, 64):
            t = Tuning(g, p, 0)
            if _lib.lib().aa_set_tuning(nstreams, nstokes, 0, ctypes.byref(t)) != 0:
                continue
            got = get_tuning(nstreams, nstokes, 0)
            out.add((got["bvp_lanes"], got["sab_pairs"]))
    set_tuning(nstreams, nstokes, 0, **orig)
    return sorted(out)


class Sample:
    """Device inputs of `ncol` seeded columns at one channel (and the vector
    Greek moments when nstokes = 3), reused across the candidates."""

    def __init__(self, torch, ncol: int, nstreams: int, nstokes: int, seed: int = 31):
        self.torch = torch
        dev = torch.device("cuda:0")
        self.ch = (550.0,)
        self.vector = nstokes == 3
        self.npol = 6 if self.vector else 1
        self.nmom = max(2 * nstreams, 16)
        self.N = nstreams
        d = S.make_columns_torch(ncol, 72, self.ch, seed=seed, col0=0, device=dev)
        self.d = d
        from . import chem_mie as AM
        self.mie = AM.MieTables.from_synthetic(S.make_mie_tables(self.ch, nmom=self.nmom, npol=self.npol))
        self.tr = torch.tensor([self.mie.query_idx(q) for q in S.TRACERS], dtype=torch.int32, device=dev)
        self.ci = torch.zeros(1, dtype=torch.int32, device=dev)
        self.rad = torch.zeros((1, ncol), dtype=torch.float64, device=dev)
        self.refl = torch.zeros_like(self.rad)
        self.q = torch.zeros_like(self.rad) if self.vector else None
        self.u = torch.zeros_like(self.rad) if self.vector else None
        self.nf = torch.zeros((1, ncol), dtype=torch.int32, device=dev)
        self.ncol = ncol

    def run(self, precision: int) -> float:
        torch = self.torch
        d = self.d
        p = lambda t: t.data_ptr() if t is not None else None
        st = torch.cuda.current_stream()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e = _lib.lib().aa_column_radiance_dev(self.mie.handle, 72, self.ncol, 1, len(S.TRACERS), p(self.tr),
                                              p(self.ci), p(d["qm"]), p(d["rh"]), self.nmom, self.npol, self.N,
                                              int(self.vector), precision, None, None, p(d["rot"]), p(d["depol"]),
                                              p(d["pe"]), p(d["he"]), p(d["te"]), p(d["albedo"]), p(d["sza"]),
                                              p(d["raa"]), p(d["vza"]), p(d["flux_factor"]), S.MISSING, p(self.rad),
                                              p(self.refl), p(self.q), p(self.u), p(self.nf), st.cuda_stream)
        _lib.check(e, "aa_column_radiance_dev")
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def outputs(self):
        return self.rad.cpu().numpy().copy(), self.nf.cpu().numpy().copy()

    def close(self):
        self.mie.close()


def tune_class(torch, nstreams: int, nstokes: int, precisions: Sequence[int], ncol: int, reps: int = 2,
               log=print) -> Dict[int, Tuple[Tuple[int, int], Dict[Tuple[int, int], float]]]:
    """Best (bvp_lanes, sab_pairs) per precision for one (streams, Stokes)
    class, with every candidate's time; raises if a candidate changes a bit of
    the results."""
    smp = Sample(torch, ncol, nstreams, nstokes)
    out = {}
    try:
        for prec in precisions:
            times = {}
            ref = None
            for g, p in shape_candidates(nstreams, nstokes):
                set_tuning(nstreams, nstokes, prec, g, p, 0)
                smp.run(prec)                              # warm-up (arena, tables)
                t = min(smp.run(prec) for _ in range(reps))
                res = smp.outputs()
                if ref is None:
                    ref = res
                elif not (np.array_equal(res[0], ref[0]) and np.array_equal(res[1], ref[1])):
                    raise RuntimeError(f"shape {(g, p)} changed the results of N={nstreams} S={nstokes}")
                times[(g, p)] = t
                log(f"N={nstreams} S={nstokes} precision={prec} bvp_lanes={g} sab_pairs={p}: {t * 1e3:.1f} ms")
            best = min(times, key=times.get)
            clear_tuning(nstreams, nstokes, prec)
            out[prec] = (best, times)
    finally:
        smp.close()
    return out


def write_table(path: str, results: Dict[Tuple[int, int], Dict[int, tuple]], ncol: int) -> None:
    lines = ["# Launch shapes of the radiance kernels on gfx950 (MI355X), chosen by aeroapps_amd/tune.py",
             f"# from the fused path on {ncol} seeded columns x 72 layers at one channel (times in ms per call).",
             "# nstreams nstokes precision bvp_lanes sab_pairs chunk_items"]
    for (N, Sx), per in sorted(results.items()):
        for prec, (best, times) in sorted(per.items()):
            alt = "  ".join(f"{g}/{p}:{t * 1e3:.1f}" for (g, p), t in sorted(times.items()))
            lines.append(f"{N} {Sx} {prec} {best[0]} {best[1]} 0   # {alt}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ncol", type=int, default=20000)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--classes", default="2,1;4,1;6,1;8,1;10,1;12,1;2,3;4,3",
                    help="semicolon-separated nstreams,nstokes pairs")
    args = ap.parse_args(argv)
    import torch
    assert torch.cuda.is_available(), "the tuner measures on the GPU"
    classes = [tuple(int(x) for x in c.split(",")) for c in args.classes.split(";") if c]
    results = {}
    for N, Sx in classes:
        results[(N, Sx)] = tune_class(torch, N, Sx, (0, 1), args.ncol,
                                      log=lambda m: print(m, file=sys.stderr, flush=True))
    write_table(args.out, results, args.ncol)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
