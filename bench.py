#!/usr/bin/env python3
"""Benchmark of the MI355X file-to-file search hot path.

Default workload: the plain form at GPU scale -- one query file of 1,000,000
intervals swept against a 1,000-file DB (log-uniform 1k-1M intervals per file,
about 1.45e8 intervals, 50 % in hotspots) on a 24-chromosome, 3.0 Gb genome.
One step = one pass of the hot path over the query: event sort + prefix sums,
merge-path partition and the sweep kernel, producing the exact 1 x 1000 u64
count and overlap-bp matrices.  Inputs are resident in HBM before timing.

Prints ONE JSON line (rank 0, N = 1).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "intervals swept/s & p-values/s on 1 MI355X, batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(db, q, nshards, nfiles, target_s=4.0, max_threads=16):
    """Reference CPU path (oracle restatement of the layered rayon sweep) on a
    bounded sample of the query; returns (record, sample counts, sample idx)."""
    import oracle
    nth = max(1, min(max_threads, os.cpu_count() or 1))
    t = time.time()
    L = oracle.LayeredDB(*db, nshards, nfiles, nthreads=nth)
    log(f"cpu baseline: layered index built in {time.time() - t:.1f}s ({L.num_layers} layers)")
    nq = len(q[1])

    def spread(k):  # k query intervals evenly strided over the sorted file: every chromosome
        return np.unique(np.linspace(0, nq - 1, k).astype(np.int64))

    probe = min(nq, 20_000)
    sel = spread(probe)
    t = time.time()
    L.query(q[0][sel], q[1][sel], q[2][sel], np.zeros(probe, np.uint32), 1, nthreads=nth)
    dt = max(time.time() - t, 1e-6)
    n_s = int(min(nq, max(probe, probe * target_s / dt)))
    sel = spread(n_s)
    n_s = len(sel)
    times, used = [], 1
    for _ in range(3):
        t = time.time()
        cnt, bins, used = L.query(q[0][sel], q[1][sel], q[2][sel], np.zeros(n_s, np.uint32), 1,
                                  nthreads=nth)
        times.append(time.time() - t)
    tmed = float(np.median(times))
    t_full = tmed * nq / n_s
    rate = (nq + len(db[1])) / t_full
    # the same scan with exact accumulators: the parity reference for counts and bp
    xc, xb = L.query_exact(q[0][sel], q[1][sel], q[2][sel], np.zeros(n_s, np.uint32), 1, nthreads=nth)
    rec = {"value": rate, "unit": "intervals/s", "cores": int(used), "kind": "port",
           "sample": (f"reference layered CPU sweep (oracle restatement, 1 query file => 1 window "
                      f"=> {used} thread as the reference schedules it) on {n_s} of the "
                      f"{nq} query intervals, evenly strided over the file, vs the full DB; "
                      f"median of 3 runs = {tmed:.3f}s, "
                      f"scaled linearly to the full query ({t_full:.2f}s)"),
           "host_threads_available": os.cpu_count()}
    return rec, (cnt, xc, xb), sel, L


def cpu_serving_baseline(L, files, F, frac=0.05, min_n=2000, max_threads=16):
    """The reference CPU path on the serving pool: one query_database call per
    request file, which is one 64-file window and so one thread.  Each
    file's time comes from max(min_n, frac * n) of its intervals, evenly
    strided over the sorted file, scaled
    by size; capacity assumes the host threads each serve one request at a
    time and scale perfectly (a bound in the CPU's favour)."""
    nth = max(1, min(max_threads, os.cpu_count() or 1))
    est = []
    for sh, s, e in files:
        n = len(s)
        k = min(n, max(min_n, int(frac * n)))
        idx = np.unique(np.linspace(0, n - 1, k).astype(np.int64))
        k = len(idx)
        t = time.time()
        L.query(sh[idx], s[idx], e[idx], np.zeros(k, np.uint32), 1, nthreads=1)
        est.append((n, (time.time() - t) * n / k))
    mean_s = float(np.mean([x[1] for x in est]))
    est.sort()
    return {"value": nth / mean_s, "unit": "requests/s", "cores": nth, "kind": "port",
            "sample": (f"reference layered CPU sweep (oracle restatement), each of the {len(files)} pool "
                       f"files timed alone on 1 thread over max({min_n}, {frac:.0%}) of its intervals, "
                       f"evenly strided, "
                       f"and scaled by size; capacity = {nth} threads / mean request time "
                       f"({mean_s * 1e3:.0f} ms), perfect scaling assumed"),
            "alone_ms_by_size": [(n, round(x * 1e3, 1)) for n, x in est]}


def cpu_batched_baseline(ix, L, batch, swept, F, target_s=4.0, max_threads=16):
    """The reference CPU path on the C4 batch: the same strided fraction of
    every one of the 256 files (so the batch keeps its four 64-file windows,
    the reference's unit of parallelism), timed on the host threads and
    scaled to the whole batch; the sample's matrix is checked bit-exact
    against the GPU's."""
    sh, s, e, qf, qc = batch
    B = len(qc)
    nth = max(1, min(max_threads, os.cpu_count() or 1))
    off = np.concatenate([[0], np.cumsum(qc)])
    rank = np.arange(len(qf), dtype=np.int64) - off[qf]

    def pick(phi):  # every (1/phi)-th interval of each sorted file, its first one included
        step = max(1, int(round(1.0 / phi)))
        return np.nonzero(rank % step == 0)[0]

    def run(sel):
        t = time.time()
        cnt, bins, used = L.query(sh[sel], s[sel], e[sel], qf[sel], B, nthreads=nth)
        return time.time() - t, cnt, used

    phi = 0.002
    dt, _, _ = run(pick(phi))
    phi = min(1.0, phi * target_s / max(dt, 1e-3))
    sel = pick(phi)
    frac = len(sel) / len(qf)
    times = []
    for _ in range(3):
        dt, cnt, used = run(sel)
        times.append(dt)
    tmed = float(np.median(times))
    t_full = tmed / frac
    q = ix.query(sh[sel], s[sel], e[sel], qf[sel], B)
    gc, gb = ix.sweep(q)
    xc, xb = L.query_exact(sh[sel], s[sel], e[sel], qf[sel], B, nthreads=nth)
    return {"value": swept / t_full, "unit": "intervals swept/s", "cores": int(used), "kind": "port",
            "sample": (f"reference layered CPU sweep (oracle restatement) on {frac:.2%} of every query "
                       f"file's intervals, evenly strided ({len(sel)} of {len(qf)}; 256 files = 4 windows "
                       f"of 64 => {used} threads as the reference schedules them); median of 3 runs "
                       f"= {tmed:.3f}s, scaled linearly to the whole batch ({t_full:.1f}s)"),
            # the reference's u32 counts are the exact counts modulo 2^32
            "sample_counts_bit_exact": bool(((gc & np.uint64(0xFFFFFFFF)) == cnt.astype(np.uint64)).all()
                                            and (gc.astype(np.int64) == xc).all()),
            "sample_bp_bit_exact": bool((gb.astype(np.int64) == xb).all())}


def profiled_traffic(workload):
    """HBM bytes per sweep launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over this same bench command (scaled by the plain copy kernel's
    reading), as committed in profiles/; None when absent."""
    path = os.path.join(ROOT, "profiles", f"sweep_{workload}_plain.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def stats_leg(ix, qq, cfg, F, steps):
    """Moment tables of the whole DB (FFT on the device), then p-values per
    nonzero cell of the query; returns timings and throughputs."""
    ix.set_genome(np.array([x for _, x in cfg["genome"]], np.uint64))
    ix.profile(True)
    for name in ("moments_fft", "moments_eval"):
        ix.profile_read(name)
    t = time.perf_counter()
    rows, exact = ix.build_moments()
    t_mom = time.perf_counter() - t
    k_fft = ix.profile_read("moments_fft")
    k_eval = ix.profile_read("moments_eval")
    so, sl, _, _ = ix.download_moments()
    nb = np.array([(x + 99) // 100 for _, x in cfg["genome"]], np.int64)
    S = len(nb)
    bins = int(sum(nb[k % S] for k in np.nonzero(so >= 0)[0]))
    log(f"moments: {rows} signals ({exact} exact), {bins} bins in {t_mom:.2f}s")
    # query groups need the genome: recreate the query handle
    sh, s, e, _ = cfg["query"]
    q2 = ix.query(sh, s, e)
    cells = F
    dc, db_ = ix.alloc(8 * cells), ix.alloc(8 * cells)
    dp, dl, dfl = ix.alloc(8 * cells), ix.alloc(8 * cells), ix.alloc(cells)
    ix.sweep_async(q2, dc, db_)
    ix.pvalues_async(q2, dc, db_, dp, dl, dfl)
    ix.sync()
    ix.profile_read("pvalues")
    t = time.perf_counter()
    for _ in range(steps):
        ix.sweep_async(q2, dc, db_)
        ix.pvalues_async(q2, dc, db_, dp, dl, dfl)
    ix.sync()
    ms_step = (time.perf_counter() - t) * 1e3 / steps
    n_pv, ms_pv = ix.profile_read("pvalues")
    ix.profile(False)
    flags = dfl.to_numpy(np.uint8, cells)
    emitted = int((flags & 1).sum())
    stats_leg.outputs = {"q": q2, "p": dp.to_numpy(np.float64, cells),
                         "llr": dl.to_numpy(np.float64, cells), "flags": flags,
                         "bp": db_.to_numpy(np.uint64, cells)}
    pv_ms = ms_pv / max(n_pv, 1)
    return {"moment_signals": rows, "moment_signals_exact": exact, "moment_bins": bins,
            "moments_build_s": t_mom, "moments_bins_per_s": bins / t_mom,
            "moments_fft_ms": k_fft[1], "moments_eval_ms": k_eval[1],
            "pvalue_cells_emitted": emitted, "pvalue_kernel_ms": pv_ms,
            "pvalues_per_s_kernel": emitted / (pv_ms * 1e-3),
            "sweep_plus_pvalues_ms_per_step": ms_step,
            "pvalues_per_s_end_to_end": emitted / (ms_step * 1e-3)}


C3_WIDTHS = [1, 2, 3, 4, 6, 10, 16, 25, 40, 63, 100, 158, 251, 398, 631, 1000]
FP32_PEAK_TFLOPS = 157.3


def c3_leg(ix, cfg, F):
    """Precision form of the moments (C3): the whole DB's 24,000 tracks at
    the 16 widths through the FP32 lag-limited FFT route, timed per kernel;
    the per-width error against the exact-integer oracle comes from the CPU
    leg."""
    ix.set_genome(np.array([x for _, x in cfg["genome"]], np.uint64))
    ix.profile(True)
    regions = ("m32_coverage", "m32_fft", "m32_eval", "m32_exact")
    for name in regions:
        ix.profile_read(name)
    t = time.perf_counter()
    m, v, info = ix.moments_widths(C3_WIDTHS)
    t_all = time.perf_counter() - t
    k = {name: ix.profile_read(name) for name in regions}
    ix.profile(False)
    c3_leg.outputs = (m, v)
    bins = info["bins"]
    fft_ms = k["m32_fft"][1]
    # 4096-point complex FFTs: two forward per block pair's blocks and one
    # inverse per pair, 5 N log2 N flops each
    blocks = 0
    for c, (_, size) in enumerate(cfg["genome"]):
        n = (size + 99) // 100
        blocks += int(np.count_nonzero(~np.isnan(m[:, c, 0]))) * ((n + 3071) // 3072)
    rows = int(np.count_nonzero(~np.isnan(m[:, :, 0])))
    flops = blocks * 1.5 * 5 * 4096 * 12
    alg_bytes = 4 * bins
    return {"widths_bins": C3_WIDTHS, "tracks": rows, "bins": bins, "wall_s": t_all,
            "kernels_ms": {n: x[1] for n, x in k.items()},
            "fft_flops": flops, "fft_tflops": flops / (fft_ms * 1e-3) / 1e12,
            "roofline": {"bound": "fp32", "frac_fp32": flops / (fft_ms * 1e-3) / (FP32_PEAK_TFLOPS * 1e12),
                         "alg_bytes": alg_bytes,
                         "frac_hbm": alg_bytes / (fft_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)},
            "exact_widths": info["exact_widths"], "inexact_rows": info["inexact_rows"],
            "route": "FP32 4096-point block FFTs (hop 3072) for lags < 1024, FP32 two-sum accumulation, "
                     "FP64 variance step; (signal, width) pairs whose FP32 error bound is not small go "
                     "through exact window sums (exact_widths)"}


def cpu_moments_and_precision(ix, cfg, F, stats_out):
    """Checks of the moment / p-value path against the oracle (FP64, exact
    integer moments) and the oracle's reference-method moment time.
    Moments: exact tables of sampled (file, chromosome) tracks vs the device
    tables, at all stored block sizes and at the 16 widths 1..1000 bins.
    p-values: every emitted cell of the query vs the oracle's FP64 evaluation
    of compute_analytic_stats on the device tables."""
    import oracle
    sh, s, e, fid = cfg["db"]
    genome = cfg["genome"]
    S = len(genome)
    so, sl, gmean, gvar = ix.download_moments()
    sample = [(f, c) for c in (S - 4, S - 3) for f in range(0, F, max(1, F // 8))][:16]
    worst = {"mean_all": 0.0, "var_all": 0.0, "mean_c3": 0.0, "var_c3": 0.0}
    tracks = []
    fp32_err = {"mean": [0.0] * len(C3_WIDTHS), "var": [0.0] * len(C3_WIDTHS)}
    order = np.lexsort((sh, fid))
    fs, ss = fid[order], sh[order]
    for f, c in sample:
        lo = np.searchsorted(fs, f, "left")
        hi = np.searchsorted(fs, f, "right")
        sel = order[lo:hi][ss[lo:hi] == c]
        if len(sel) == 0 or so[f * S + c] < 0:
            continue
        nb = (genome[c][1] + 99) // 100
        ls, xm, xv = oracle.chrom_moments_exact(s[sel], e[sel], nb)
        o, n = so[f * S + c], sl[f * S + c]
        gm, gv = gmean[o:o + n], gvar[o:o + n]
        rel = lambda a, b: float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))
        worst["mean_all"] = max(worst["mean_all"], rel(gm, xm))
        worst["var_all"] = max(worst["var_all"], rel(gv, xv))
        idx = [oracle.compact_index(L) for L in C3_WIDTHS if L < nb]
        worst["mean_c3"] = max(worst["mean_c3"], rel(gm[idx], xm[idx]))
        worst["var_c3"] = max(worst["var_c3"], rel(gv[idx], xv[idx]))
        tracks.append((s[sel], e[sel], nb, ls))
        if hasattr(c3_leg, "outputs"):  # the FP32 route vs the exact oracle at the 16 widths
            ok = [i for i, L in enumerate(C3_WIDTHS) if L < nb]
            xm3, xv3, _, _ = oracle.moments_exact(oracle.depth_exact(s[sel], e[sel], nb),
                                                  np.array([C3_WIDTHS[i] for i in ok], np.uint32))
            m3, v3 = c3_leg.outputs
            for jj, i in enumerate(ok):
                em = abs(m3[f, c, i] - xm3[jj]) / max(abs(xm3[jj]), 1e-300)
                ev = abs(v3[f, c, i] - xv3[jj]) / max(abs(xv3[jj]), 1e-300)
                fp32_err["mean"][i] = max(fp32_err["mean"][i], float(em))
                fp32_err["var"][i] = max(fp32_err["var"][i], float(ev))
    # the reference method on the CPU (f32 depth signal, FP64 FFT moments),
    # one track per host thread as the reference's par_iter over files runs
    # them (the oracle's C calls release the GIL)
    import concurrent.futures as cf

    def ref_track(tr):
        oracle.moments_reference(oracle.depth_signal_f32(tr[0], tr[1], tr[2]), tr[3])

    nth = max(1, min(16, os.cpu_count() or 1, len(tracks)))
    with cf.ThreadPoolExecutor(max_workers=nth) as ex:
        t = time.perf_counter()
        list(ex.map(ref_track, tracks))
        t_ref = time.perf_counter() - t
    bins_ref = sum(tr[2] for tr in tracks)
    # p-values of the query: oracle FP64 evaluation on the device tables
    qsh, qs, qe, _ = cfg["query"]
    groups = {}
    for c in range(S):
        m = qsh == c
        if m.any():
            groups[c] = oracle.group_lengths(qe[m] - qs[m])
    dmax, checked, t_pv = 0.0, 0, 0.0
    for f in range(F):
        if not stats_out["flags"][f] & 1:
            continue
        tabs, parts = [], [[], [], [], [], []]
        for c, g4 in groups.items():
            o = so[f * S + c]
            slot = -1
            if o >= 0:
                slot = len(tabs)
                tabs.append(((genome[c][1] + 99) // 100, gmean[o:o + sl[f * S + c]],
                             gvar[o:o + sl[f * S + c]]))
            for a, v in zip(parts, g4 + (np.full(len(g4[0]), slot, np.int32),)):
                a.append(v)
        args_ = (tuple(np.concatenate(a) for a in parts), None, tabs, float(stats_out["bp"][f]) / 100.0)
        t0 = time.perf_counter()
        res = oracle.analytic_stats(*args_)
        t_pv += time.perf_counter() - t0
        if res is None:
            continue
        ml = lambda p: np.inf if p <= 0 else -np.log10(p)
        a, b = ml(stats_out["p"][f]), ml(res[0])
        if np.isfinite(a) and np.isfinite(b):
            dmax = max(dmax, abs(a - b))
        checked += 1
    out_fp32 = {"widths_bins": C3_WIDTHS, "mean_rel_err_max_per_width": fp32_err["mean"],
                "var_rel_err_max_per_width": fp32_err["var"], "tracks_checked": len(tracks)} \
        if hasattr(c3_leg, "outputs") else None
    return {"moments_rel_err_max": worst, "moment_signals_checked": len(sample),
            "fp32_route_vs_exact": out_fp32,
            "c3_widths_bins": C3_WIDTHS, "pvalue_cells_checked": checked,
            "pvalue_abs_err_mlog10p_max": dmax,
            "pvalues_cpu": {"value": checked / max(t_pv, 1e-9), "unit": "p-values/s", "cores": 1,
                            "kind": "port",
                            "sample": f"the oracle's FP64 compute_analytic_stats on all {checked} emitted "
                                      f"cells of the query, one thread (timed per call, including the "
                                      f"ctypes marshalling of each call)"},
            "moments_reference_method_cpu": {"bins_per_s": bins_ref / max(t_ref, 1e-9),
                                             "cores": nth, "kind": "port",
                                             "sample": f"{len(tracks)} (file, chromosome) tracks, "
                                                       f"one per thread"}}


def batched_leg(ix, cfg, F, nshards, D):
    """256 query files (log-uniform 100-2M intervals, 30 % hotspot) against the
    same DB.  Plain: one launch per file, each file on its own best path, then
    dense-only (the plain merge-path method).  Batched: the autotuner picks the
    knobs on the batch, then the batch runs as one captured HIP graph."""
    from riggle_amd import gen
    t = time.time()
    g = cfg["genome"]
    hot = gen.hotspots(g, 0xC40C1E000 + 2)
    qc = gen.log_uniform_counts(256, 100, 2e6, 0xC40C1E000 + 4, stream=0xC4C4)
    sh, s, e = gen.intervals(g, int(qc.sum()), 0xC40C1E000 + 4, 2, 1000.0, 1.2,
                             hotspot_frac=0.3, hot=hot)
    sh, s, e, qf = gen.sort_files(sh, s, e, np.repeat(np.arange(256, dtype=np.uint32), qc))
    off = np.concatenate([[0], np.cumsum(qc)])
    log(f"batched: generated 256 files, {int(qc.sum())} intervals in {time.time() - t:.1f}s")
    B = 256
    d1, d2 = ix.alloc(8 * F), ix.alloc(8 * F)
    singles = [ix.query(sh[off[i]:off[i + 1]], s[off[i]:off[i + 1]], e[off[i]:off[i + 1]])
               for i in range(B)]

    def plain_loop():
        ix.sweep_async(singles[0], d1, d2)
        ix.sync()
        t = time.perf_counter()
        for i in range(B):
            ix.sweep_async(singles[i], d1, d2)
        ix.sync()
        return (time.perf_counter() - t) * 1e3

    ms_plain = plain_loop()
    ix.tune("sweep.method", 1)
    ms_plain_dense = plain_loop()
    ix.tune("sweep.method", 0)
    qb = ix.query(sh, s, e, qf, B)
    dc, db_ = ix.alloc(8 * B * F), ix.alloc(8 * B * F)
    # the autotuner sees a different batch of the same laws (another seed),
    # so the timed batch is not the one the knobs were picked on
    tc = gen.log_uniform_counts(256, 100, 2e6, 0xC40C1E000 + 40, stream=0xC4C4)
    tsh, ts, te = gen.intervals(g, int(tc.sum()), 0xC40C1E000 + 40, 2, 1000.0, 1.2,
                                hotspot_frac=0.3, hot=hot)
    qt = ix.query(tsh, ts, te, np.repeat(np.arange(256, dtype=np.uint32), tc), B)
    t = time.perf_counter()
    best, knobs, report = ix.autotune(qt, dc, db_, reps=2)
    t_tune = time.perf_counter() - t
    qt.close()
    log(f"batched: autotune on a second batch {t_tune:.1f}s best {best:.2f} ms {knobs}")
    ix.sweep_async(qb, dc, db_)
    ix.sync()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        ix.sweep_async(qb, dc, db_)
    ix.sync()
    ms_eager = (time.perf_counter() - t) * 1e3 / reps
    plan = qb.info()  # dense / sparse split under the tuning the graph uses
    graph = ix.graph(qb, dc, db_)
    graph.launch()
    ix.sync()
    t = time.perf_counter()
    for _ in range(reps):
        graph.launch()
    ix.sync()
    ms_graph = (time.perf_counter() - t) * 1e3 / reps
    cb = dc.to_numpy(np.uint64, (B, F))
    rows_ok = True
    for i in (0, 1, B // 2, B - 1):
        ix.sweep_async(singles[i], d1, d2)
        rows_ok &= bool((d1.to_numpy(np.uint64, F) == cb[i]).all())
    graph.close()
    for k in ("sweep.sparse_max", "sweep.items", "sweep.chunk", "sweep.blocks_per_cu"):
        ix.tune(k, 0)
    swept = B * D + int(qc.sum())
    batched_leg.batch = (sh, s, e, qf, qc)
    # DB bytes: each dense file streams the DB once (g = 1: files do not share
    # staged DB tiles); sparse files read only the layered-table windows they
    # visit, which this figure leaves out
    nq = int(qc.sum())
    bytes_batch = 12 * D * plan["dense_files"] + 8 * nq + 16 * B * F
    bytes_effective = 12 * D * B + 8 * nq + 16 * B * F
    t_s = ms_graph * 1e-3
    return {"query_files": B, "query_intervals": nq, "intervals_swept": swept,
            "dense_files": plan["dense_files"], "dense_intervals": plan["dense_intervals"],
            "db_stream_group_g": 1,
            "roofline": {"bytes_batch": bytes_batch, "frac": bytes_batch / (t_s * HBM_PEAK_GBS * 1e9),
                         "bytes_effective": bytes_effective,
                         "frac_effective": bytes_effective / (t_s * HBM_PEAK_GBS * 1e9)},
            "batched_graph_ms": ms_graph, "batched_eager_ms": ms_eager,
            "plain_ms_256_launches": ms_plain, "plain_dense_ms_256_launches": ms_plain_dense,
            "batched_swept_per_s": swept / (ms_graph * 1e-3),
            "plain_swept_per_s": swept / (ms_plain * 1e-3),
            "plain_dense_swept_per_s": swept / (ms_plain_dense * 1e-3),
            "batched_over_plain": ms_plain / ms_graph,
            "batched_over_plain_dense": ms_plain_dense / ms_graph,
            "autotuned": knobs, "autotune_s": t_tune, "rows_match_single_launches": rows_ok,
            "autotune_note": "0 keeps the automatic choice: no candidate below beat it",
            "autotune_candidates": report,
            "autotune_batch": "a second 256-file batch of the same laws (seed + 40); the timed batch "
                              "was not seen by the tuner"}


def pcie_leg(ix, cfg, F, D, steps):
    """The plain form with the query handed over in host memory each step:
    query upload (validation, pinned staging, H2D copy, plan) + sweep + the
    two result rows copied back.  Reported beside `value`, never as it."""
    qsh, qs, qe, _ = cfg["query"]
    dc, db_ = ix.alloc(8 * F), ix.alloc(8 * F)
    q = ix.query(qsh, qs, qe)
    for _ in range(2):
        q.reset(qsh, qs, qe)
        ix.sweep_async(q, dc, db_)
        dc.to_numpy(np.uint64, F)
    t = time.perf_counter()
    for _ in range(steps):
        q.reset(qsh, qs, qe)
        ix.sweep_async(q, dc, db_)
        dc.to_numpy(np.uint64, F)
        db_.to_numpy(np.uint64, F)
    ms = (time.perf_counter() - t) * 1e3 / steps
    t = time.perf_counter()
    for _ in range(steps):
        q2 = ix.query(qsh, qs, qe)
        ix.sweep_async(q2, dc, db_)
        dc.to_numpy(np.uint64, F)
        db_.to_numpy(np.uint64, F)
        q2.close()
    ms_new = (time.perf_counter() - t) * 1e3 / steps
    return {"ms_per_step": ms, "swept_per_s": (len(qs) + D) / (ms * 1e-3), "steps": steps,
            "region": "host query arrays -> chk_query_reset on a reused handle (host validation, "
                      "pinned staging, H2D copy, plan) -> sweep -> counts and bp rows copied to the host",
            "ms_per_step_new_handle": ms_new}


def c5_leg(args, parity):
    """The serving configuration: a 10,000-file DB of about 2.15e9 intervals
    resident in HBM.  Plain 1M-interval query steps, the Poisson serving mix
    (>= 1,000 requests of 1k-5M intervals, 40 % hotspot), one end-to-end
    request from a 5M-interval BED file (parse over host threads, sweep,
    JSON), and the reference path's CPU serving baseline on a bounded
    sample."""
    import riggle_amd
    from riggle_amd import gen
    t = time.time()
    cfg = gen.config("c5")
    sh, s, e, fid = cfg["db"]
    S, F, D = len(cfg["genome"]), cfg["nfiles"], len(cfg["db"][1])
    log(f"c5: generated |D|={D} F={F} in {time.time() - t:.1f}s")
    t = time.time()
    ix = riggle_amd.Index(sh, s, e, fid, S, F)
    t_index = time.time() - t
    info = ix.info()
    log(f"c5: index built in {t_index:.1f}s ({info['device_bytes'] / 1e9:.1f} GB)")
    qsh, qs, qe, _ = cfg["query"]
    qq = ix.query(qsh, qs, qe)
    dc, db_ = ix.alloc(8 * F), ix.alloc(8 * F)
    for _ in range(5):
        ix.sweep_async(qq, dc, db_)
    ix.sync()
    ix.profile(True, only=("sweep",))
    ix.profile_read("sweep")
    steps = 20
    t0 = time.perf_counter()
    for _ in range(steps):
        ix.sweep_async(qq, dc, db_)
    ix.sync()
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    n_k, ms_k = ix.profile_read("sweep")
    ix.profile(False)
    k_ms = ms_k / max(n_k, 1)
    alg = 12 * D + 8 * len(qs) + 16 * F
    rc_ms, rc_bytes = ix.calibrate_read(3)
    out = {"db_files": F, "db_intervals": D, "index_device_bytes": info["device_bytes"],
           "read_ceiling_gbs": rc_bytes / (rc_ms * 1e-3) / 1e9,
           "index_build_s": t_index,
           "plain": {"query_intervals": len(qs), "ms_per_step": ms_step,
                     "swept_per_s": (len(qs) + D) / (ms_step * 1e-3), "sweep_kernel_ms": k_ms,
                     "roofline_frac": alg / (k_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)}}
    log(f"c5 plain: {ms_step:.2f} ms/step, kernel {k_ms:.2f} ms")
    serving = serving_leg(ix, cfg, F, D, args.c5_requests, pool=48)
    out["serving"] = serving
    log(f"c5 serving: p50 {serving['p50_ms']:.2f} ms p99 {serving['p99_ms']:.2f} ms "
        f"capacity {serving['capacity_req_per_s']:.1f} req/s")
    # one request end to end from text: a 5M-interval BED file
    import tempfile
    import pandas as pd
    names = [n for n, _ in cfg["genome"]]
    ix.set_names(names, [f"db_{k:05d}.bed" for k in range(F)])
    big = gen.intervals(cfg["genome"], 5_000_000, 0xC40C1E000 + 5, 77, 1000.0, 1.2,
                        hotspot_frac=0.4, hot=cfg["hot"])
    tmp = tempfile.mkdtemp(prefix="chk_bench_")
    bed = os.path.join(tmp, "query_5m.bed")
    pd.DataFrame({"c": np.array(names)[big[0]], "s": big[1], "e": big[2]}).to_csv(
        bed, sep="\t", header=False, index=False)
    js = os.path.join(tmp, "out.json")
    e2e = None
    for rep in range(2):  # the second pass is reported (the first warms page caches)
        t0 = time.perf_counter()
        q, rows = ix.query_files([bed])
        t1 = time.perf_counter()
        ix.sweep_async(q, dc, db_)
        c = dc.to_numpy(np.uint64, F)
        b = db_.to_numpy(np.uint64, F)
        t2 = time.perf_counter()
        riggle_amd.write_results(js, ["query_5m.bed"], ix.names()[1], c.reshape(1, F), b.reshape(1, F))
        t3 = time.perf_counter()
        e2e = {"query_intervals": 5_000_000, "bed_bytes": os.path.getsize(bed),
               "parse_and_upload_ms": (t1 - t0) * 1e3, "sweep_and_copy_ms": (t2 - t1) * 1e3,
               "json_ms": (t3 - t2) * 1e3, "json_bytes": os.path.getsize(js),
               "records": int((c > 0).sum()), "total_ms": (t3 - t0) * 1e3,
               "host_threads": os.cpu_count()}
        q.close()
    # the parsed file gives the same counts as the generated arrays
    qg = ix.query(*big)
    ix.sweep_async(qg, dc, db_)
    parity["c5_parsed_file_rows"] = bool((dc.to_numpy(np.uint64, F) == c).all())
    out["end_to_end_request"] = e2e
    log(f"c5 end to end (5M-interval BED): {e2e}")
    for f in (bed, js):
        os.remove(f)
    os.rmdir(tmp)
    if "cpu" in args.legs_set:
        import oracle
        t = time.time()
        L = oracle.LayeredDB(sh, s, e, fid, S, F, nthreads=16)
        log(f"c5 cpu: layered index built in {time.time() - t:.1f}s")
        cs = cpu_serving_baseline(L, serving_leg.files[:16], F, frac=0.01, min_n=1000)
        out["serving"]["cpu_baseline"] = cs
        out["serving"]["capacity_vs_cpu"] = serving["capacity_req_per_s"] / cs["value"]
        log(f"c5 cpu serving baseline {cs['value']:.3f} req/s")
        del L
    ix.close()
    return out


def serving_leg(ix, cfg, F, D, n_requests, seed=0xC40C1E000 + 5, pool=32):
    """Serving form against the resident DB: query files of log-uniform 1k-5M
    intervals (40 % hotspot) arrive as a Poisson stream at 65 % of the
    measured capacity; latency is submission to completion."""
    import threading
    import riggle_amd
    from riggle_amd import gen
    t = time.time()
    g = cfg["genome"]
    hot = cfg.get("hot") or gen.hotspots(g, 0xC40C1E000 + 2)
    sizes = gen.log_uniform_counts(pool, 1e3, 5e6, seed, stream=0xC5C5)
    files = [gen.sort_files(*gen.intervals(g, int(n), seed, 100 + i, 1000.0, 1.2, hotspot_frac=0.4,
                                           hot=hot))
             for i, n in enumerate(sizes)]
    log(f"serving: {pool} distinct query files ({int(sizes.sum())} intervals) in "
        f"{time.time() - t:.1f}s")
    serving_leg.files = files
    srv = riggle_amd.Server(ix, slots=3, max_batch_files=16)
    # any batch of this pool (at most 16 files) allocates nothing
    srv.reserve(int(np.sort(sizes)[-16:].sum()))
    # each distinct file alone (latency by size, after one pass over the pool
    # has grown the slots' buffers), then capacity: closed loop over the pool,
    # twice (the first pass warms up)
    for f in files:
        srv.wait(srv.submit(*f))
    alone = []
    for f in files:
        r = srv.wait(srv.submit(*f))
        alone.append((len(f[1]), round(r["latency_ms"], 3)))
    alone.sort()
    for rep in range(2):
        t = time.perf_counter()
        tk = [srv.submit(*f) for f in files]
        for x in tk:
            srv.wait(x)
        dt = time.perf_counter() - t
    cap = pool / dt
    rate = 0.65 * cap
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    pick = rng.integers(0, pool, n_requests)
    srv.reset_stats()
    tickets = []
    done = threading.Event()

    def waiter():
        k = 0
        while k < n_requests:
            if k < len(tickets):
                srv.wait(tickets[k])
                k += 1
            else:
                time.sleep(0.0002)
        done.set()

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    due = t0
    for i in range(n_requests):
        due += gaps[i]
        while True:
            now = time.perf_counter()
            if now >= due:
                break
            time.sleep(min(due - now, 0.001))
        tickets.append(srv.submit(*files[pick[i]]))
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    swept = sum(len(files[i][1]) + D for i in pick)
    return {"requests": n_requests, "distinct_files": pool, "sizes": "log-uniform 1k-5M intervals",
            "capacity_req_per_s": cap, "arrival_rate_per_s": rate, "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "files_per_batch": st["files_per_batch"], "swept_per_s": st["swept_per_s"],
            "offered_swept_per_s": swept / wall, "slots": 3, "max_batch_files": 16,
            "alone_ms_by_size": alone}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c5"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tune", action="append", default=[],
                    help="launch-shape knob key=value (see chk_tune), may repeat")
    ap.add_argument("--skip-stats", action="store_true", help="skip moments + p-value legs")
    ap.add_argument("--skip-batched", action="store_true", help="skip the batched 256-file leg")
    ap.add_argument("--serve-requests", type=int, default=1000,
                    help="requests of the serving leg (0 skips it)")
    ap.add_argument("--c5-requests", type=int, default=1000,
                    help="requests of the serving leg on the 10,000-file C5 DB")
    ap.add_argument("--legs", default=None,
                    help="comma list of legs to run (plain,pcie,stats,c3,batched,serving,cpu,c5); the "
                         "plain leg gives the JSON line's value")
    ap.add_argument("--calibrate-bytes", type=int, default=0,
                    help="also run the plain copy kernel over this many bytes (counter scaling)")
    args = ap.parse_args()
    if args.legs is None:  # the serving DB is too large for the CPU and moment legs
        args.legs = ("plain,serving" if args.config == "c5"
                     else "plain,pcie,stats,c3,batched,serving,cpu,c5")
    legs = set(args.legs.split(","))
    if args.config != "c2":
        legs.discard("c5")
    if args.no_cpu:
        legs.discard("cpu")
    if args.skip_stats:
        legs.discard("stats")
        legs.discard("c3")
    if args.skip_batched:
        legs.discard("batched")
    if args.serve_requests <= 0:
        legs.discard("serving")
    args.legs_set = legs
    parity = {}
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import riggle_amd
    from riggle_amd import gen

    t = time.time()
    cfg = gen.config(args.config)
    sh, s, e, fid = cfg["db"]
    qsh, qs, qe, _ = cfg["query"]
    nshards, F = len(cfg["genome"]), cfg["nfiles"]
    D, Q = len(s), len(qs)
    log(f"generated {args.config}: |D|={D} |Q|={Q} F={F} in {time.time() - t:.1f}s")

    t = time.time()
    ix = riggle_amd.Index(sh, s, e, fid, nshards, F)
    t_index = time.time() - t
    log(f"index built in {t_index:.2f}s: {ix.info()}")
    for kv in args.tune:
        k, v = kv.split("=")
        ix.tune(k, int(v))
    qq = ix.query(qsh, qs, qe)
    dc, db_ = ix.alloc(8 * F), ix.alloc(8 * F)

    steps = args.steps if "plain" in legs else 1
    warm = args.warmup if "plain" in legs else 0
    for _ in range(warm):
        ix.sweep_async(qq, dc, db_)
    ix.sync()
    log("warmup done")
    # timed region: HIP events around the dominant kernel only (each recorded
    # region costs the stream a few microseconds)
    ix.profile(True, only=("sweep", "sparse"))
    for name in ("sweep", "sparse"):
        ix.profile_read(name)
    ix.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ix.sweep_async(qq, dc, db_)
    ix.sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / steps
    kern = {name: ix.profile_read(name) for name in ("sweep",)}
    # per-stage breakdown from a separate pass of the same steps, every region recorded
    ix.profile(True)
    for _ in range(steps):
        ix.sweep_async(qq, dc, db_)
    kern.update({name: ix.profile_read(name) for name in ("qprep", "partition")})
    ix.profile(False)
    counts = dc.to_numpy(np.uint64, F)
    bp = db_.to_numpy(np.uint64, F)
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step; kernels {kern}")
    calib = None
    if args.calibrate_bytes > 0:
        ix.profile(True)
        ix.calibrate_copy(args.calibrate_bytes)
        n_c, ms_c = ix.profile_read("calib_copy")
        ix.profile(False)
        calib = {"bytes": args.calibrate_bytes, "ms": ms_c,
                 "gbs_rw": 2 * args.calibrate_bytes / (ms_c * 1e-3) / 1e9}
        log(f"calibration copy: {calib}")

    rc_ms, rc_bytes = ix.calibrate_read(3)
    read_ceiling = {"bytes": rc_bytes, "ms": rc_ms, "gbs": rc_bytes / (rc_ms * 1e-3) / 1e9,
                    "kernel": "calib_read_kernel: 16-byte loads of the dense streams only, best of "
                              "3 runs at each of 3 grid sizes"}
    log(f"read ceiling of the dense streams: {read_ceiling['gbs']:.0f} GB/s")
    swept = (Q + D) / (ms_step * 1e-3)
    n_sweep, ms_sweep = kern["sweep"]
    sweep_ms = ms_sweep / max(n_sweep, 1)
    alg_bytes = 12 * D + 8 * Q + 16 * F
    achieved = alg_bytes / (sweep_ms * 1e-3) / 1e9

    rec = {
        "metric": METRIC, "value": swept, "unit": "intervals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32 compare, u64/int64 accumulate", "data": "synthetic",
        "config": {"workload": f"{args.config} plain: 1 query file vs {F}-file DB",
                   "db_intervals": D, "query_intervals": Q, "db_files": F,
                   "timed_region": "query resident in HBM: event sort + prefix sums, merge-path "
                                   "partition, sweep kernel, exact u64 count and bp rows "
                                   "(PCIe-inclusive rate under pcie_inclusive)",
                   "genome_bp": int(sum(x for _, x in cfg["genome"])), "chromosomes": nshards},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args.config),
                     "kernel": "sweep_kernel", "kernel_ms": sweep_ms,
                     "alg_bytes_per_launch": alg_bytes,
                     "read_ceiling_gbs": read_ceiling["gbs"],
                     "frac_of_read_ceiling": achieved / read_ceiling["gbs"]},
        "read_ceiling": read_ceiling,
        "kernels_ms_per_step": {k: (v[1] / max(v[0], 1)) for k, v in kern.items()},
        "index_build_s": t_index,
        "tuning": args.tune,
        "result_check": {"total_pairs": int(counts.sum()), "total_bp": int(bp.sum())},
    }
    if calib:
        rec["calibration_copy"] = calib

    if "pcie" in legs:
        rec["pcie_inclusive"] = pcie_leg(ix, cfg, F, D, 10)
        log(f"pcie-inclusive: {rec['pcie_inclusive']['ms_per_step']:.3f} ms/step")
    if "stats" in legs:
        rec["stats"] = stats_leg(ix, qq, cfg, F, args.steps)
    if "c3" in legs:
        rec.setdefault("stats", {})["c3_fp32_route"] = c3_leg(ix, cfg, F)
        log(f"c3 fp32 route: {rec['stats']['c3_fp32_route']['kernels_ms']}")
    if "batched" in legs and args.config == "c2":
        rec["batched"] = batched_leg(ix, cfg, F, nshards, D)
        parity["batched_rows_match_single_launches"] = rec["batched"]["rows_match_single_launches"]
    if "serving" in legs and args.config in ("c2", "c5"):
        rec["serving"] = serving_leg(ix, cfg, F, D, args.serve_requests)
        log(f"serving: {rec['serving']}")

    if "cpu" in legs:
        rec_cpu, (cnt_cpu, xc, xb), sel, L = cpu_baseline((sh, s, e, fid), (qsh, qs, qe), nshards, F)
        # parity on the same sample: GPU counts and bp must equal the reference
        # scan's (exact accumulators; its u32 counts agree modulo 2^32)
        qs_ = ix.query(qsh[sel], qs[sel], qe[sel])
        gc, gb = ix.sweep(qs_)
        rec_cpu["sample_counts_bit_exact"] = bool((gc[0].astype(np.int64) == xc[0]).all() and
                                                  ((gc[0] & np.uint64(0xFFFFFFFF)) ==
                                                   cnt_cpu[0].astype(np.uint64)).all())
        rec_cpu["sample_bp_bit_exact"] = bool((gb[0].astype(np.int64) == xb[0]).all())
        parity["plain_sample_counts"] = rec_cpu["sample_counts_bit_exact"]
        parity["plain_sample_bp"] = rec_cpu["sample_bp_bit_exact"]
        if hasattr(stats_leg, "outputs"):
            t = time.time()
            rec_cpu["moments_and_pvalues"] = mp = cpu_moments_and_precision(ix, cfg, F, stats_leg.outputs)
            parity["moments_within_1e-6"] = bool(max(mp["moments_rel_err_max"].values()) <= 1e-6)
            parity["pvalues_within_1e-3_mlog10p"] = bool(mp["pvalue_abs_err_mlog10p_max"] <= 1e-3)
            if mp["fp32_route_vs_exact"]:
                e3 = mp["fp32_route_vs_exact"]
                parity["c3_fp32_route_within_1e-6"] = bool(
                    max(e3["mean_rel_err_max_per_width"] + e3["var_rel_err_max_per_width"]) <= 1e-6)
            log(f"cpu: moment / p-value checks in {time.time() - t:.1f}s: "
                f"{rec_cpu['moments_and_pvalues']['moments_rel_err_max']}")
        rec["cpu_baseline"] = rec_cpu
        rec["speedup_vs_cpu"] = swept / rec_cpu["value"]
        if "batched" in rec and hasattr(batched_leg, "batch"):
            t = time.time()
            cb = cpu_batched_baseline(ix, L, batched_leg.batch, rec["batched"]["intervals_swept"], F)
            rec["batched"]["cpu_baseline"] = cb
            parity["batched_sample_counts"] = cb["sample_counts_bit_exact"]
            parity["batched_sample_bp"] = cb["sample_bp_bit_exact"]
            rec["batched"]["speedup_vs_cpu"] = rec["batched"]["batched_swept_per_s"] / cb["value"]
            log(f"cpu batched baseline {cb['value']:.3e} intervals swept/s on {cb['cores']} threads "
                f"in {time.time() - t:.1f}s; sample parity {cb['sample_counts_bit_exact']}")
        if "serving" in rec and hasattr(serving_leg, "files"):
            t = time.time()
            cs = cpu_serving_baseline(L, serving_leg.files, F)
            rec["serving"]["cpu_baseline"] = cs
            rec["serving"]["capacity_vs_cpu"] = rec["serving"]["capacity_req_per_s"] / cs["value"]
            log(f"cpu serving baseline {cs['value']:.2f} req/s on {cs['cores']} threads in "
                f"{time.time() - t:.1f}s")
        log(f"cpu baseline {rec_cpu['value']:.3e} intervals/s; sample parity "
            f"{rec_cpu['sample_counts_bit_exact']}")
    if "c5" in legs:
        ix.close()
        rec["c5"] = c5_leg(args, parity)
    rec["parity"] = parity
    ok = all(parity.values())
    if not ok:  # a wrong result publishes no rate
        rec["status"] = "parity_failed"
        rec["value_if_correct"] = rec["value"]
        rec["value"] = None
    print(json.dumps(rec), flush=True)
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
