"""Benchmark of the aerosol-optics + discrete-ordinate radiance path on one GPU.

Workload (default): 500,000 columns x 72 layers, 4 channels (354, 388, 550,
870 nm), 12 streams (6 per hemisphere), 64 Legendre moments, 15 tracers,
Lambertian albedo 0.05 -- seeded synthetic inputs generated on the device.  A
step is one pass of the fused device path (Chem_Mie optics -> layer optics ->
radiance) over every column, with inputs resident in HBM.

Prints one JSON line (metric, value in columns/s, ms_per_step, the dominant
kernel's roofline figures measured with HIP events on its launch stream, and
the host-CPU reference timed on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from aeroapps_amd import synthetic as S  # noqa: E402

METRIC = "columns/sec at 72 layers/12 streams, mixed vs FP64 GPU; p99 granule latency"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
PROF_NAMES = ["mie", "check", "optics", "geom", "item", "layer", "bvp", "brdf", "sabdab"]
# HBM bytes per item-term (read, written) of the radiance kernels at N=6
# scalar, keyed (kernel, N, S, mixed), from FETCH_SIZE/WRITE_SIZE counter runs
# scaled by a copy-kernel calibration
# (profiles/hbm_traffic_40k_columns_4ch_12streams.md)
TRAFFIC_PER_ITEM_TERM = {("bvp", 6, 1, False): (153.2e3, 10.9e3), ("bvp", 6, 1, True): (153.2e3, 10.9e3),
                         ("layer", 6, 1, False): (55.2e3, 84.1e3), ("layer", 6, 1, True): (34.6e3, 86.7e3)}


# ----------------------------------------------------------------- algorithmic work
def flops_per_term(N, S_, L, nmoments, M):
    """Algorithmic FP64 flops of one (column, channel) at Fourier term M:
    eigensystem (F_b), boundary-value solve (F_c) and source (F_d).  F_b is
    returned without its SAB/DAB contraction term 4 s n^2 Lm per layer (that
    contraction runs in k_sabdab, counted by sabdab_per_term), i.e. the work
    k_layer itself does: omega-Greek x PI products, EIGENMAT = DAB.SAB and the
    nonsymmetric eigensolve with vectors.  F_c counts the structurally non-zero
    work of the pivoted factorisation (one panel of 3n rows x (4n columns +
    RHS) per layer, 2n x (2n + RHS) for the bottom layer) and of the adjoint
    solve U^T w = c with the running w . b' (per eliminated column: w_j, its
    product with b'_j, and the update of the accumulators of the row's
    remaining columns, per Stokes output); F_d the source functional c, d
    (per layer S (2n + 2) + 1).  An FMA counts as 2 flops."""
    n = S_ * N
    n2 = 2 * n
    Lm = nmoments - M + 1
    Fb = L * (4 * N * S_ ** 3 * Lm + 2 * n ** 3 + 25 * n ** 3)
    mid = sum((3 * n - 1 - i) * (1 + 2 * (4 * n - i)) for i in range(n2))
    last = sum((2 * n - 1 - i) * (1 + 2 * (2 * n - i)) for i in range(n2))
    adj_mid = S_ * sum(4 + 2 * (2 * n2 - 1 - i) for i in range(n2))
    adj_last = S_ * sum(4 + 2 * (n2 - 1 - i) for i in range(n2))
    Fc = (L - 1) * (mid + adj_mid) + last + adj_last
    Fd = L * (S_ * (n2 + 2) + 1)
    return Fb, Fc, Fd


def sabdab_per_term(N, S_, L, nmoments, M, elem):
    """Algorithmic flops and HBM bytes of the SAB/DAB contraction of one
    (column, channel) at Fourier term M: per layer the 2n^2 SAB and DAB entries
    are each a dot product over the Lm * S^2 non-zero omega-Greek coefficients
    (l >= M); bytes are those coefficients read plus the entries written, in
    elements of `elem` bytes."""
    n = S_ * N
    k = (nmoments - M + 1) * S_ * S_
    return L * 2 * (2 * n * n) * k, L * elem * (2 * n * n + k)


def bytes_per_column(km, nq, nch, vector):
    """Compulsory HBM bytes of the fused path per column (inputs + outputs)."""
    b = 4 * km * nq + 4 * km + 8 * km * nch + 24 * (km + 1) + 24 + 8 * nch
    if vector:
        b += 8 * nch
    b += 8 * nch * (2 + 2 * int(vector))
    return b


# ----------------------------------------------------------------- CPU baseline
def _cpu_worker(args):
    (c0, nc, wl, kind, modis) = args
    import oracle
    from oracle import chem_mie as CM
    drv = (oracle.ref_modis() if modis else oracle.ref()) if kind == "reference" else oracle.port()
    ch = wl["channels"]
    t0 = time.perf_counter()
    cols = S.make_columns(nc, wl["km"], ch, seed=7, col0=c0)
    tabs = S.make_mie_tables(ch, nmom=wl["nmom"], npol=wl["npol"])
    gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    tau, ssa, g, pm, rc = CM.get_aop_vector(tabs, ch, tabs[0].lambda_m, S.TRACERS, cols.qm, cols.rh, wl["nmom"],
                                            wl["npol"])
    f = lambda a: np.asarray(a, dtype=np.float64)
    if modis:
        kw, par = S.make_modis_kernels(nc, len(ch), seed=7, col0=c0)
        drv.scalar_landmodis(wl["km"], cols.channels, wl["nstreams"], wl["nmom"], cols.rot, cols.depol, f(tau),
                             f(ssa), f(g), f(pm), cols.pe, cols.he, cols.te, kw, par, cols.sza, cols.raa, cols.vza)
    else:
        drv.scalar_lambert(wl["km"], cols.channels, wl["nstreams"], wl["nmom"], cols.rot, cols.depol, f(tau), f(ssa),
                           f(g), f(pm), cols.pe, cols.he, cols.te, cols.albedo, cols.sza, cols.raa, cols.vza)
    return time.perf_counter() - t0, gen


def host_cpu():
    """(model name, logical CPUs of the host, CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_baseline(wl, ncols, procs, modis=False):
    import oracle
    kind = "reference" if os.path.exists(oracle.REF_SO) else "port"
    if kind == "port" and (modis or not os.path.exists(oracle.PORT_SO)):
        return None
    per = max(1, ncols // procs)
    jobs = [(i * per, per, wl, kind, modis) for i in range(procs)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    busy = max(r[0] for r in res)
    ncol = per * procs
    model, total, avail = host_cpu()
    return {
        "value": ncol / busy,
        "unit": "columns/s",
        "cores": procs,
        "kind": kind,
        "cpu_model": model,
        "host_logical_cpus": total,
        "cpus_available_to_process": avail,
        "sample": (f"{ncol} columns ({per} per process, {procs} processes, one core each) of the same workload: "
                   f"float32 Chem_Mie restatement + "
                   + ("reference VLIDORT 2.8.3 core (flang -O3) under the oracle's statement-by-statement "
                      "restatement of the AeroApps driver set-up, bit-identical to the reference's compiled "
                      "drivers on the test fixtures"
                      if kind == "reference" else "C++ restatement")
                   + f", scalar {'MODIS land-surface' if modis else 'Lambertian'}, {len(wl['channels'])} channels, "
                   f"N={wl['nstreams']}, nmom={wl['nmom']}; slowest process {busy:.1f} s, pool wall {wall:.1f} s; "
                   f"processes = the CPUs this process may run on, at most 16 (the host CPU share of one GPU "
                   f"on the measuring box)"),
    }


# ----------------------------------------------------------------- GPU run
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="scalar_500k_4ch")
    ap.add_argument("--ncol", type=int, default=0, help="override the column count")
    ap.add_argument("--fp64-only", action="store_true", help="skip the mixed-precision run")
    ap.add_argument("--cpu-cols", type=int, default=1024)
    ap.add_argument("--cpu-procs", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-granules", type=int, default=40,
                    help="granules of the serving-loop phase (0 skips it)")
    ap.add_argument("--serving-load", type=float, default=0.7, help="offered load as a fraction of peak")
    ap.add_argument("--surface", choices=("lambert", "modis"), default="lambert",
                    help="surface model: Lambertian albedo or the MODIS land BRDF (scalar workloads)")
    ap.add_argument("--host-cols", type=int, default=20000,
                    help="columns of the host-pointer (PCIe-inclusive) drop-in phase (0 skips it)")
    ap.add_argument("--no-other-configs", dest="other_configs", action="store_false",
                    help="skip the vector (200k x 2 ch) and uneven (300k) configurations of the default run")
    ap.add_argument("--work-gib", type=float, default=64.0,
                    help="work-arena budget of the throughput runs in GiB (one stream; 0 keeps the library default)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the path has no cross-GPU step (every column is independent); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    wl = dict(S.WORKLOADS[args.workload])
    ncol = args.ncol or wl["ncol"]
    km, nch, nq = wl["km"], len(wl["channels"]), len(S.TRACERS)
    N, nmom, npol, vector = wl["nstreams"], wl["nmom"], wl["npol"], wl["vector"]
    uneven = wl.get("uneven_layers")
    if N is None:   # per-column stream counts (uneven batch)
        N = max(wl["stream_choices"])

    # host-CPU baseline first, before this process touches the GPU (fork-safe)
    cpu = None
    if not args.no_cpu:
        procs = args.cpu_procs or min(16, host_cpu()[2])
        cpu = cpu_baseline(wl, args.cpu_cols, procs, modis=args.surface == "modis")

    import torch
    from aeroapps_amd import _lib
    from aeroapps_amd import chem_mie as AM
    assert torch.cuda.is_available(), "bench needs a GPU"
    dev = torch.device("cuda:0")
    L = _lib.lib()

    # inputs, generated on the device in slices
    t0 = time.perf_counter()
    parts = []
    step = 100_000
    for c0 in range(0, ncol, step):
        parts.append(S.make_columns_torch(min(step, ncol - c0), km, wl["channels"], seed=7, col0=c0, device=dev,
                                          uneven_layers=uneven))
    cat = lambda k, dim: torch.cat([p[k] for p in parts], dim=dim).contiguous()
    qm = cat("qm", 0)            # [nobs, nq, km]  == Fortran qm(km,nq,nobs)
    rh = cat("rh", 0)            # [nobs, km]
    rot = cat("rot", 1)          # [nch, nobs, km]
    he = cat("he", 0)            # [nobs, km+1]
    pe = cat("pe", 0)
    te = cat("te", 0)
    albedo = cat("albedo", 1)    # [nch, nobs]
    sza, vza, raa = cat("sza", 0), cat("vza", 0), cat("raa", 0)
    flux = cat("flux_factor", 0)  # [nobs, nch]
    depol = parts[0]["depol"]
    nl_h = torch.cat([p["nlayers"] for p in parts]).cpu().numpy()
    d_nl = torch.cat([p["nlayers"] for p in parts]).contiguous() if uneven else None
    if wl["nstreams"] is None:
        nstr_h = S.stream_choice(ncol, wl["stream_choices"], seed=7)
        d_nstr = torch.tensor(nstr_h, dtype=torch.int32, device=dev)
    else:
        nstr_h = np.full(ncol, N, dtype=np.int32)
        d_nstr = None
    del parts
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0

    tabs = S.make_mie_tables(wl["channels"], nmom=nmom, npol=npol)
    mie = AM.MieTables.from_synthetic(tabs)
    tr_idx = torch.tensor([mie.query_idx(n) for n in S.TRACERS], dtype=torch.int32, device=dev)
    ch_idx = torch.arange(nch, dtype=torch.int32, device=dev)
    rad = torch.empty((nch, ncol), dtype=torch.float64, device=dev)
    refl = torch.empty_like(rad)
    qo = torch.empty_like(rad) if vector else None
    uo = torch.empty_like(rad) if vector else None
    nf = torch.empty((nch, ncol), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    ptr = lambda t: t.data_ptr() if t is not None else None

    modis = args.surface == "modis"
    if modis:
        assert not vector, "the MODIS land surface is scalar"
        kw_h, par_h = S.make_modis_kernels(ncol, nch, seed=7)
        d_kw = torch.as_tensor(kw_h, dtype=torch.float64, device=dev).contiguous()    # (3,nch,nobs) Fortran
        d_par = torch.as_tensor(par_h, dtype=torch.float64, device=dev).contiguous()  # (2,nch,nobs) Fortran
        br = torch.empty_like(rad)
        del kw_h, par_h

    def run_step(precision):
        if modis:
            e = L.aa_column_radiance_modis_dev(mie.handle, km, ncol, nch, nq, ptr(tr_idx), ptr(ch_idx), ptr(qm),
                                               ptr(rh), nmom, N, precision, ptr(d_nl), ptr(d_nstr), ptr(rot),
                                               ptr(depol), ptr(pe), ptr(he), ptr(te), ptr(d_kw), ptr(d_par), ptr(sza),
                                               ptr(raa), ptr(vza), S.MISSING, ptr(rad), ptr(refl), ptr(br), ptr(nf),
                                               stream.cuda_stream)
            _lib.check(e, "aa_column_radiance_modis_dev")
            return
        e = L.aa_column_radiance_dev(mie.handle, km, ncol, nch, nq, ptr(tr_idx), ptr(ch_idx), ptr(qm), ptr(rh), nmom,
                                     npol, N, int(vector), precision, ptr(d_nl), ptr(d_nstr), ptr(rot), ptr(depol),
                                     ptr(pe),
                                     ptr(he), ptr(te), ptr(albedo), ptr(sza), ptr(raa), ptr(vza), ptr(flux),
                                     S.MISSING, ptr(rad), ptr(refl), ptr(qo), ptr(uo), ptr(nf), stream.cuda_stream)
        _lib.check(e, "aa_column_radiance_dev")

    import ctypes
    nmoments = min(2 * N - 1, nmom)
    Sx = 3 if vector else 1

    def timed(precision):
        for i in range(args.warmup):
            run_step(precision)
            torch.cuda.synchronize()
            print(f"precision {precision}: warmup step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        L.aa_profile_start()
        t0 = time.perf_counter()
        for i in range(args.steps):
            run_step(precision)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        nk = len(PROF_NAMES)
        msa = (ctypes.c_double * nk)()
        lna = (ctypes.c_longlong * nk)()
        _lib.check(L.aa_profile_read(msa, lna, nk), "aa_profile_read")
        keep = [k for k in range(nk) if k < 7 or (k == 7 and modis) or (k == 8 and lna[k] > 0)]
        kms = {PROF_NAMES[k]: msa[k] for k in keep}
        kln = {PROF_NAMES[k]: lna[k] for k in keep}
        ms_per_step = el * 1000.0 / args.steps
        nfh = nf.cpu().numpy().ravel().copy()
        radh = rad.cpu().numpy().ravel().copy()
        prods = {"reflectance": refl.cpu().numpy().ravel().copy()}
        if vector:
            prods["q"] = qo.cpu().numpy().ravel().copy()
            prods["u"] = uo.cpu().numpy().ravel().copy()
        print(f"precision {precision}: {args.steps} timed steps done, {ms_per_step:.1f} ms/step", file=sys.stderr,
              flush=True)
        return dict(el=el, kms=kms, kln=kln, ms_per_step=ms_per_step, value=ncol / (ms_per_step / 1000.0), nf=nfh,
                    rad=radh, prods=prods)

    if args.work_gib > 0:
        _lib.check(L.aa_set_arena_bytes(int(args.work_gib * (1 << 30)), 0), "aa_set_arena_bytes")
    r64 = timed(0)
    rmx = timed(1) if not args.fp64_only else None
    main_run = rmx if rmx is not None else r64
    if args.work_gib > 0:
        # the serving slots and the host-pointer phase run with the default budgets
        _lib.check(L.aa_release_arenas(), "aa_release_arenas")
        _lib.check(L.aa_set_arena_bytes(0, 0), "aa_set_arena_bytes")

    # algorithmic work of the timed region from the Fourier-term counts
    nfh = main_run["nf"]
    ok = nfh > 0
    fb = fc = fd = 0.0
    sab_f = sab_b4 = sab_b8 = 0.0
    # items are (channel, column) in nf's order; each with its column's N and L
    n_it = np.tile(nstr_h, nch)[ok]
    l_it = np.tile(nl_h, nch)[ok]
    combos, cnts = np.unique(np.stack([n_it, l_it, nfh[ok]]), axis=1, return_counts=True)
    for (n_c, l_c, terms), cnt in zip(combos.T, cnts):
        for M in range(int(terms)):
            b_, c_, d_ = flops_per_term(int(n_c), Sx, int(l_c), min(2 * int(n_c) - 1, nmom), M)
            fb += cnt * b_; fc += cnt * c_; fd += cnt * d_
            f_, b4_ = sabdab_per_term(int(n_c), Sx, int(l_c), min(2 * int(n_c) - 1, nmom), M, 4)
            sab_f += cnt * f_; sab_b4 += cnt * b4_
            sab_b8 += cnt * sabdab_per_term(int(n_c), Sx, int(l_c), min(2 * int(n_c) - 1, nmom), M, 8)[1]
    fb *= args.steps; fc *= args.steps; fd *= args.steps
    sab_f *= args.steps; sab_b4 *= args.steps; sab_b8 *= args.steps

    def mfma_use(run, f32):
        """Matrix-core use of the SAB/DAB contraction in one timed run."""
        ms = run["kms"].get("sabdab", 0.0)
        if ms <= 0.0:
            return None
        peak = PEAK_FP32_MFMA_TFLOPS if f32 else PEAK_FP64_TFLOPS
        tf = sab_f / (ms / 1000.0) / 1e12
        gbs = (sab_b4 if f32 else sab_b8) / (ms / 1000.0) / 1e9
        return {"kernel": f"k_sabdab<{N},{Sx},{'float' if f32 else 'double'}>", "ms": round(ms, 3),
                "achieved": round(tf, 3), "peak": peak, "unit": "TFLOP/s", "frac": round(tf / peak, 4),
                "algorithmic_GBs": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}
    kms, kln = main_run["kms"], main_run["kln"]
    item_terms = float(nfh[ok].sum()) * args.steps
    # compulsory HBM bytes of the path per (item, Fourier term): a column's
    # inputs and outputs spread over its channels' executed terms
    compulsory_item_term = bytes_per_column(km, nq, nch, vector) * ncol * args.steps / max(item_terms, 1.0)

    def roofline(kname):
        """FP64-VALU roofline of one radiance kernel over the timed region:
        algorithmic flops (k_layer: F_b without the SAB/DAB contraction, which
        k_sabdab performs; k_bvp: structural F_c + F_d) over the kernel's total
        HIP-event time; HBM traffic per launch from the counter constants."""
        ms = kms[kname]
        launches = max(1, kln[kname])
        flops = fb if kname == "layer" else fc + fd
        ach = flops / (ms / 1000.0) / 1e12
        r = {"bound": "fp64_valu", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
             "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": None,
             "kernel": (f"k_{kname}<{N},{Sx}>" if wl["nstreams"] is not None
                        else f"k_{kname}<{min(wl['stream_choices'])}..{N},{Sx}>"),
             "kernel_avg_ms": round(ms / launches, 4), "kernel_launches": int(launches),
             "flops": ("F_b without the SAB/DAB term" if kname == "layer" else "F_c + F_d (structural)"),
             "note": "FP64 vector ALU work (no MFMA); peak is the MI355X FP64 vector rate"}
        key = (kname, N, Sx, rmx is not None)
        if key in TRAFFIC_PER_ITEM_TERM and wl["nstreams"] is not None and not uneven and not modis:
            rd, wr = TRAFFIC_PER_ITEM_TERM[key]
            r["traffic"] = round((rd + wr) * item_terms / launches, 0)
            r["traffic_GBs"] = round((rd + wr) * item_terms / (ms / 1000.0) / 1e9, 1)
            r["hbm_frac"] = round(r["traffic_GBs"] / PEAK_HBM_GBS, 4)
            r["traffic_vs_compulsory"] = round((rd + wr) / compulsory_item_term, 1)
        return r

    dom = max(("layer", "bvp"), key=lambda k: kms[k])
    el = main_run["el"]
    value = main_run["value"]
    path_bytes = bytes_per_column(km, nq, nch, vector) * ncol * args.steps
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if rmx is not None else "f64",
        "data": "synthetic (seeded, generated on device)",
        "config": {"workload": f"{args.workload}: {ncol} columns x "
                               f"{f'{uneven[0]}-{uneven[1]}' if uneven else km} layers, {nch} channels "
                               f"{list(wl['channels'])} nm, "
                               f"{2 * N if wl['nstreams'] is not None else '/'.join(str(2 * c) for c in wl['stream_choices'])}"
                               f" streams, {nmom} moments, {nq} tracers, "
                               f"{'vector' if vector else 'scalar'} "
                               f"{'MODIS land BRDF (Lambertian + Ross-thick + Li-sparse)' if modis else 'Lambertian'}",
                   "precision": "mixed (FP32 Chem_Mie optics, FP32 phase-matrix contraction on MFMA, FP32 source "
                                "assembly from 8 streams on; FP64 eigensystem and boundary-value elimination; the "
                                "FP64 line is FP64 throughout with an FP64 MFMA contraction)"
                                if rmx is not None else "fp64",
                   "path": ("aa_column_radiance_modis_dev" if modis else "aa_column_radiance_dev") +
                           " (Chem_Mie optics + VLIDORT radiance, device-resident)",
                   "work_arena_gib": args.work_gib if args.work_gib > 0 else 24.0},
        "roofline": roofline(dom),
        "roofline_other": roofline("bvp" if dom == "layer" else "layer"),
        "hbm_path": {"bytes_per_column": bytes_per_column(km, nq, nch, vector),
                     "achieved_GBs": round(path_bytes / el / 1e9, 3),
                     "frac": round(path_bytes / el / 1e9 / PEAK_HBM_GBS, 6)},
        "kernel_ms": {k: round(v, 3) for k, v in kms.items()},
        "sabdab_mfma": {"mixed": mfma_use(rmx, True) if rmx is not None else None, "fp64": mfma_use(r64, False)},
        "fourier_terms_mean": round(float(nfh[ok].mean()), 3),
        "fp64": {"value": round(r64["value"], 1), "ms_per_step": round(r64["ms_per_step"], 3),
                 "kernel_ms": {k: round(v, 3) for k, v in r64["kms"].items()}},
        "cpu_baseline": cpu,
        "gen_s": round(gen_s, 2),
    }
    if rmx is not None:
        same = (rmx["nf"] == r64["nf"]) & ok
        rel = np.abs(rmx["rad"] - r64["rad"])[same] / np.abs(r64["rad"][same])
        out["mixed_vs_fp64"] = {"speedup": round(rmx["value"] / r64["value"], 3),
                                "max_rel_err_radiance": float(rel.max()) if rel.size else None,
                                "p99_rel_err_radiance": float(np.percentile(rel, 99)) if rel.size else None,
                                "fourier_count_mismatch_frac": float(1 - same.sum() / ok.sum())}
        # error of each product against the FP64 path (elements with the same
        # Fourier-term count): relative for radiance and reflectance, Q and U
        # relative to |I|
        per = {"radiance": rel}
        i64 = np.abs(r64["rad"][same])
        for k, v in rmx["prods"].items():
            d = np.abs(v[same] - r64["prods"][k][same])
            per[k] = d / (np.abs(r64["prods"][k][same]) if k == "reflectance" else i64)
        out["mixed_vs_fp64"]["per_product"] = {
            k: {"max": float(v.max()) if v.size else None, "p99": float(np.percentile(v, 99)) if v.size else None,
                "relative_to": "itself" if k in ("radiance", "reflectance") else "|I|"} for k, v in per.items()}
    if cpu:
        out["gpu_vs_cpu"] = round(value / cpu["value"], 1)
        out["fp64_vs_cpu"] = round(r64["value"] / cpu["value"], 1)

    # serving loop: Poisson granules at a fraction of the measured peak
    if args.serving_granules > 0 and not vector and wl["nstreams"] is not None and not uneven and not modis:
        from aeroapps_amd import serving as SV
        del qm, rh, rot, he, pe, te, albedo, sza, vza, raa, flux, rad, refl, nf
        torch.cuda.empty_cache()
        pool = SV.ColumnPool(torch, 200_000, km, wl["channels"], dev)
        grans = SV.make_granules(args.serving_granules, pool.ncol, wl["channels"])
        mean_cols = float(np.mean([g.ncol for g in grans]))
        rate = args.serving_load * value / mean_cols
        arr = SV.poisson_arrivals(len(grans), rate)
        srv = SV.Server(torch, mie, pool, N, nmom, slots=4, hbm_budget=200e9, arena_bytes=28e9,
                        precision=1 if rmx is not None else 0)
        st = srv.serve(grans, arr)
        out["serving"] = {"granules": st["granules"], "columns": st["columns"],
                          "offered_load": args.serving_load, "arrival_rate_per_s": round(rate, 4),
                          "precision": "mixed" if rmx is not None else "fp64",
                          "p50_latency_s": round(st["p50_s"], 3), "p99_latency_s": round(st["p99_s"], 3),
                          "throughput_cols_per_s": round(st["throughput_cols_per_s"], 1), "slots": st["slots"],
                          "granule_columns": "log-uniform 1000-200000, 1-4 channels, 72 layers, 12 streams"}
    # drop-in host-pointer path (PCIe-inclusive): getAOPvector + VLIDORT_Scalar_Lambert
    if args.host_cols > 0 and not vector and wl["nstreams"] is not None and not uneven and not modis:
        from aeroapps_amd import vlidort as AV
        hc = min(args.host_cols, ncol)
        cols_h = S.make_columns(hc, km, wl["channels"], seed=7)
        chs = np.asarray(wl["channels"], dtype=np.float32)
        torch.cuda.synchronize()
        times = []
        for _ in range(2):   # the first pass also grows the library's host-call device block
            t0 = time.perf_counter()
            tau_h, ssa_h, g_h, pm_h, rc_a = AM.get_aop_vector(mie, chs, S.TRACERS, 0, cols_h.qm.T, cols_h.rh.T,
                                                              nmom, npol)
            t1 = time.perf_counter()
            rad_h, refl_h, rc_l = AV.scalar_lambert(cols_h.channels, N, 0, cols_h.rot.transpose(2, 1, 0),
                                                    cols_h.depol, tau_h, ssa_h, g_h, pm_h, cols_h.pe.T, cols_h.he.T,
                                                    cols_h.te.T, cols_h.albedo.T, cols_h.sza, cols_h.raa, cols_h.vza,
                                                    S.MISSING, 0)
            t2 = time.perf_counter()
            times.append((t0, t1, t2))
        (f0, _, f2), (t0, t1, t2) = times
        out["host_path"] = {"columns": hc, "getaop_s": round(t1 - t0, 3), "lambert_s": round(t2 - t1, 3),
                            "columns_per_s": round(hc / (t2 - t0), 1),
                            "first_call_columns_per_s": round(hc / (f2 - f0), 1), "rc": [int(rc_a), int(rc_l)],
                            "note": "host numpy arrays in, host arrays out (FP64 optics widened on the host, "
                                    "PCIe copies inside the calls); not the headline value"}
    # the other device-resident configurations, each in a child process of
    # its own (one bench line; this process first gives its device memory back)
    if args.other_configs and args.workload == "scalar_500k_4ch" and not args.ncol:
        _lib.check(L.aa_release_arenas(), "aa_release_arenas")
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        out["other_configs"] = {w: other_config(w) for w in ("vector_200k_2ch", "uneven_300k_1ch")}
    print(json.dumps(out))


def other_config(workload):
    """One warm-up and one timed step of `workload` in both precisions, from a
    child run of this script; the rates and step times of its bench line."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", workload, "--steps", "1", "--warmup", "1",
           "--no-cpu", "--serving-granules", "0", "--host-cols", "0", "--no-other-configs"]
    print(f"other configuration {workload}: starting", file=sys.stderr, flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)   # progress lines go straight to stderr
    if p.returncode != 0:
        return {"error": f"exit {p.returncode}"}
    line = json.loads(p.stdout.strip().splitlines()[-1])
    r = {"workload": line["config"]["workload"], "columns_per_s_mixed": line["value"],
         "ms_per_step_mixed": line["ms_per_step"], "columns_per_s_fp64": line["fp64"]["value"],
         "ms_per_step_fp64": line["fp64"]["ms_per_step"], "steps": 1, "warmup": 1}
    if "mixed_vs_fp64" in line:
        r["mixed_vs_fp64"] = {k: line["mixed_vs_fp64"][k] for k in ("speedup", "max_rel_err_radiance",
                                                                    "fourier_count_mismatch_frac")}
    r["kernel_ms_mixed"] = line.get("kernel_ms")
    return r


if __name__ == "__main__":
    main()
