"""CPU oracle for Chuckle's file-to-file search path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package, and only as the
checker or the CPU baseline.  The product (``riggle_amd``) never imports it.

This module is a thin ctypes/numpy layer over ``liboracle.so`` (built from the
C files beside it by ``make -C oracle``); every function documents the
reference file:line it restates.  See ``oracle.h`` for the list.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CHK_ORACLE_LIB selects another build of the same sources (the sanitizer
# build of tests/sanitize); by default the library beside this file
_LIB_PATH = os.environ.get("CHK_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

CHROMS24 = [f"chr{i}" for i in range(1, 23)] + ["chrX", "chrY"]  # src/bench.rs:15-19

# src/fourier/genome.rs:7-34 (hg38 chromosome sizes; default genome table)
HG38 = [
    ("chr1", 248_956_422), ("chr2", 242_193_529), ("chr3", 198_295_559),
    ("chr4", 190_214_555), ("chr5", 181_538_259), ("chr6", 170_805_979),
    ("chr7", 159_345_973), ("chr8", 145_138_636), ("chr9", 138_394_717),
    ("chr10", 133_797_422), ("chr11", 135_086_622), ("chr12", 133_275_309),
    ("chr13", 114_364_328), ("chr14", 107_043_718), ("chr15", 101_991_189),
    ("chr16", 90_338_345), ("chr17", 83_257_441), ("chr18", 80_373_285),
    ("chr19", 58_617_616), ("chr20", 64_444_167), ("chr21", 46_709_983),
    ("chr22", 50_818_468), ("chrX", 156_040_895), ("chrY", 57_227_415),
]
BIN_SIZE = 100  # src/fourier/genome.rs:3


def build() -> None:
    """Compile liboracle.so from the C sources beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _p(dt):
    return np.ctypeslib.ndpointer(dtype=dt, flags="C_CONTIGUOUS")


def _declare(L):
    u32p, u64p, i64p, f32p, f64p, u8p, i32p = (
        _p(np.uint32), _p(np.uint64), _p(np.int64), _p(np.float32), _p(np.float64),
        _p(np.uint8), _p(np.int32))
    sz = C.c_size_t
    L.orc_chacha_block.argtypes = [u32p, C.c_uint64, C.c_uint64, C.c_int, u32p]
    L.orc_rng_seed_u64.argtypes = [C.c_void_p, C.c_uint64]
    L.orc_rng_next_u32.argtypes = [C.c_void_p]
    L.orc_rng_next_u32.restype = C.c_uint32
    L.orc_rng_next_u64.argtypes = [C.c_void_p]
    L.orc_rng_next_u64.restype = C.c_uint64
    L.orc_rng_range_u32_incl.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
    L.orc_rng_range_u32_incl.restype = C.c_uint32
    L.orc_generate_bed.argtypes = [C.c_uint64, sz, C.c_uint32, C.c_uint32, C.c_uint32, u8p, u32p, u32p]
    L.orc_layer_for_size.argtypes = [C.c_uint32] * 3
    L.orc_layer_for_size.restype = C.c_uint32
    L.orc_layer_max_size.argtypes = [C.c_uint32] * 3
    L.orc_layer_max_size.restype = C.c_uint32
    L.orc_tile_size.argtypes = [C.c_uint32] * 3
    L.orc_tile_size.restype = C.c_uint32
    L.orc_build_jump_table.argtypes = [u32p, sz, C.c_uint32, C.c_void_p]
    L.orc_build_jump_table.restype = sz
    L.orc_layered_build.argtypes = [sz, u32p, u32p, u32p, u32p, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.c_int]
    L.orc_layered_build.restype = C.c_void_p
    L.orc_layered_free.argtypes = [C.c_void_p]
    L.orc_layered_num_layers.argtypes = [C.c_void_p]
    L.orc_layered_num_layers.restype = C.c_uint32
    L.orc_layered_query.argtypes = [C.c_void_p, sz, u32p, u32p, u32p, u32p, C.c_uint32,
                                    u32p, f32p, C.c_int]
    L.orc_layered_query.restype = C.c_int
    L.orc_layered_query_exact.argtypes = [C.c_void_p, sz, u32p, u32p, u32p, u32p, C.c_uint32,
                                          i64p, i64p, C.c_int]
    L.orc_layered_query_exact.restype = C.c_int
    L.orc_sweep_bruteforce.argtypes = [sz, u32p, u32p, u32p, u32p, C.c_uint32, sz, u32p, u32p,
                                       u32p, u32p, C.c_uint32, i64p, i64p]
    L.orc_sweep_ranks.argtypes = [sz, u32p, u32p, u32p, u32p, C.c_uint32, sz, u32p, u32p, u32p,
                                  u32p, C.c_uint32, C.c_uint32, i64p, i64p]
    L.orc_compact_index.argtypes = [sz]
    L.orc_compact_index.restype = sz
    L.orc_compact_ls.argtypes = [C.c_uint32, C.c_void_p]
    L.orc_compact_ls.restype = sz
    L.orc_depth_signal_f32.argtypes = [sz, u32p, u32p, C.c_uint32, f32p]
    L.orc_depth_exact.argtypes = [sz, u32p, u32p, C.c_uint32, i64p]
    L.orc_autocorr_fft.argtypes = [f64p, sz, f64p]
    L.orc_moments_reference.argtypes = [f32p, C.c_uint32, u32p, sz, f64p, f64p]
    L.orc_moments_exact.argtypes = [i64p, C.c_uint32, u32p, sz, f64p, f64p, i64p, f64p]
    L.orc_moments_exact_fft.argtypes = [i64p, C.c_uint32, u32p, sz, f64p, f64p]
    L.orc_moments_exact_fft.restype = C.c_int
    L.orc_erfc_as.argtypes = [C.c_double]
    L.orc_erfc_as.restype = C.c_double
    L.orc_nb_llr.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.orc_nb_llr.restype = C.c_int
    L.orc_poisson_llr.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.orc_poisson_llr.restype = C.c_int
    L.orc_group_lengths.argtypes = [sz, u32p, u32p, f64p, f64p, u64p]
    L.orc_group_lengths.restype = sz
    L.orc_analytic_stats.argtypes = [sz, u32p, f64p, f64p, u64p, i32p, u32p, u64p, u64p, f64p,
                                     f64p, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.orc_analytic_stats.restype = C.c_int


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ----------------------------------------------------------------- generator

def chacha_block(key, ctr, stream=0, rounds=20):
    out = np.zeros(16, np.uint32)
    lib().orc_chacha_block(_u32(key), ctr, stream, rounds, out)
    return out


def stdrng_u32(seed: int, n: int):
    """First n u32 outputs of rand 0.8.5 StdRng::seed_from_u64(seed)."""
    buf = (C.c_uint8 * 128)()
    lib().orc_rng_seed_u64(C.addressof(buf), seed)
    return np.array([lib().orc_rng_next_u32(C.addressof(buf)) for _ in range(n)], np.uint32)


def generate_bed(seed, n, genome_size=250_000_000, min_len=100, max_len=10_000):
    """src/bench.rs:75-104 (BedGenConfig defaults :32-44).  Returns
    (chrom index into CHROMS24 as uint8, start, end)."""
    ch = np.zeros(n, np.uint8)
    s = np.zeros(n, np.uint32)
    e = np.zeros(n, np.uint32)
    lib().orc_generate_bed(seed, n, genome_size, min_len, max_len, ch, s, e)
    return ch, s, e


def write_bed(path, chrom, start, end, names=CHROMS24):
    with open(path, "w") as f:
        for c, s, e in zip(chrom.tolist(), start.tolist(), end.tolist()):
            f.write(f"{names[c]}\t{s}\t{e}\n")


# --------------------------------------------------------------- Monte Carlo

class StdRng:
    """rand 0.8.5 StdRng::seed_from_u64 (rng.c), with gen_range on i64 restated
    in exact integers: UniformInt::sample_single_inclusive draws a u64 v,
    takes v * range as 128 bits and accepts when the low 64 bits are at most
    zone = (range << clz64(range)) - 1, returning lo + the high 64 bits."""

    def __init__(self, seed):
        self._buf = (C.c_uint8 * 128)()
        lib().orc_rng_seed_u64(C.addressof(self._buf), seed)

    def next_u32(self):
        return lib().orc_rng_next_u32(C.addressof(self._buf))

    def next_u64(self):
        return lib().orc_rng_next_u64(C.addressof(self._buf))

    def range_u32(self, lo, hi):
        return lib().orc_rng_range_u32_incl(C.addressof(self._buf), lo, hi)

    def range_i64(self, lo, hi):
        rng = (hi - lo + 1) % (1 << 64)
        if rng == 0:
            v = self.next_u64()
            return v - (1 << 64) if v >= 1 << 63 else v
        zone = ((rng << (64 - rng.bit_length())) - 1) % (1 << 64)
        while True:
            m = self.next_u64() * rng
            if m % (1 << 64) <= zone:
                return lo + (m >> 64)


def hg38_sizes():
    """src/fourier/genome.rs:7-34."""
    return dict(HG38)


def montecarlo_trials(qmap, seed, milestones, batch_size=100, shuffle=False, whitelist=None):
    """src/bin/montecarlo.rs:79-163 and :232-311: the trial batches the driver
    queries.  qmap: {chrom: [(start, end), ...]} in file order.  Yields
    (trials run in the batch, [non-empty trial: [(chrom, start, end), ...]])."""
    sizes = hg38_sizes()
    max_shift = max(sizes.values())
    ms = sorted(set(milestones))
    rng = StdRng(seed)
    total, mi = 0, 0
    while total < ms[-1]:
        to_run = min(ms[mi] - total, max(1, batch_size))
        batch = []
        for _ in range(to_run):
            out = []
            shift = 0 if shuffle else rng.range_i64(-max_shift, max_shift)
            for chrom in sorted(qmap):
                if chrom not in sizes:
                    continue
                size = sizes[chrom]
                wl = (whitelist or {}).get(chrom)

                def put(a, b):
                    if wl is None:
                        out.append((chrom, a, b))
                        return
                    i = 0
                    while i < len(wl) and wl[i][1] <= a:  # partition_point(end <= a)
                        i += 1
                    while i < len(wl) and wl[i][0] < b:
                        out.append((chrom, max(wl[i][0], a), min(wl[i][1], b)))
                        i += 1

                for s, e in qmap[chrom]:
                    if shuffle:
                        ln = e - s
                        if ln == 0 or ln > size:
                            continue
                        ns = rng.range_u32(0, size - ln)
                        put(ns, ns + ln)
                    else:
                        a, b = s + shift, e + shift
                        if b <= 0 or a >= size:
                            continue
                        a, b = max(a, 0), min(b, size)
                        if a >= b:
                            continue
                        put(a, b)
            if out:
                batch.append(out)
        yield to_run, batch
        total += to_run
        while mi < len(ms) and total >= ms[mi]:
            mi += 1
        if mi >= len(ms):
            break


# --------------------------------------------------------------------- sweep

def layer_for_size(size, min_size=128, gf=4):
    return lib().orc_layer_for_size(min_size, gf, size)


def layer_max_size(layer, min_size=128, gf=4):
    return lib().orc_layer_max_size(min_size, gf, layer)


def build_jump_table(sorted_starts, tile):
    s = _u32(sorted_starts)
    n = lib().orc_build_jump_table(s, len(s), tile, None)
    out = np.zeros(n, np.uint64)
    lib().orc_build_jump_table(s, len(s), tile, out.ctypes.data)
    return out


class LayeredDB:
    """Reference CPU index: size-class layers + jump tables
    (src/tasks/build.rs:196-247, src/io/meta.rs:33-81, src/io/jump.rs:66-113)."""

    def __init__(self, shard, start, end, sid, nshards, nsources, min_size=128, gf=4,
                 nthreads=1):
        self.nsources = int(nsources)
        self.nshards = int(nshards)
        self._h = lib().orc_layered_build(len(start), _u32(shard), _u32(start), _u32(end),
                                          _u32(sid), nshards, nsources, min_size, gf, nthreads)

    @property
    def num_layers(self):
        return lib().orc_layered_num_layers(self._h)

    def query(self, shard, start, end, qsid, nq, nthreads=1):
        """query_database's parallel core: (u32 counts, f32 bins, threads used)."""
        cnt = np.zeros((nq, self.nsources), np.uint32)
        bins = np.zeros((nq, self.nsources), np.float32)
        used = lib().orc_layered_query(self._h, len(start), _u32(shard), _u32(start), _u32(end),
                                       _u32(qsid), nq, cnt, bins, nthreads)
        return cnt, bins, used

    def query_exact(self, shard, start, end, qsid, nq, nthreads=1):
        """The reference's windowed scan with exact int64 (counts, bp)."""
        cnt = np.zeros((nq, self.nsources), np.int64)
        bp = np.zeros((nq, self.nsources), np.int64)
        lib().orc_layered_query_exact(self._h, len(start), _u32(shard), _u32(start), _u32(end),
                                      _u32(qsid), nq, cnt, bp, nthreads)
        return cnt, bp

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_layered_free(self._h)
            self._h = None


def sweep_bruteforce(db, q, F, nq):
    """db/q = (shard, start, end, sid) arrays; exact int64 (counts, bp)."""
    cnt = np.zeros((nq, F), np.int64)
    bp = np.zeros((nq, F), np.int64)
    lib().orc_sweep_bruteforce(len(db[1]), *[_u32(a) for a in db], F, len(q[1]),
                               *[_u32(a) for a in q], nq, cnt, bp)
    return cnt, bp


def sweep_ranks(db, q, F, nq, nshards):
    """Pair-free exact (counts, bp) via rank identities."""
    cnt = np.zeros((nq, F), np.int64)
    bp = np.zeros((nq, F), np.int64)
    lib().orc_sweep_ranks(len(db[1]), *[_u32(a) for a in db], F, len(q[1]),
                          *[_u32(a) for a in q], nq, nshards, cnt, bp)
    return cnt, bp


# ------------------------------------------------------------------- moments

def compact_index(l):
    return lib().orc_compact_index(l)


def compact_ls(n_bins):
    n = lib().orc_compact_ls(n_bins, None)
    out = np.zeros(max(n, 1), np.uint32)
    lib().orc_compact_ls(n_bins, out.ctypes.data)
    return out[:n]


def n_bins_of(size):
    return (int(size) + BIN_SIZE - 1) // BIN_SIZE


def depth_signal_f32(start, end, n_bins):
    g = np.zeros(n_bins, np.float32)
    lib().orc_depth_signal_f32(len(start), _u32(start), _u32(end), n_bins, g)
    return g


def depth_exact(start, end, n_bins):
    G = np.zeros(n_bins, np.int64)
    lib().orc_depth_exact(len(start), _u32(start), _u32(end), n_bins, G)
    return G


def autocorr_fft(x):
    x = np.ascontiguousarray(x, np.float64)
    r = np.zeros(len(x), np.float64)
    lib().orc_autocorr_fft(x, len(x), r)
    return r


def moments_reference(g, ls):
    ls = _u32(ls)
    m = np.zeros(len(ls), np.float64)
    v = np.zeros(len(ls), np.float64)
    lib().orc_moments_reference(np.ascontiguousarray(g, np.float32), len(g), ls, len(ls), m, v)
    return m, v


def moments_exact(G, ls):
    ls = _u32(ls)
    m = np.zeros(len(ls), np.float64)
    v = np.zeros(len(ls), np.float64)
    s1 = np.zeros(len(ls), np.int64)
    s2 = np.zeros(len(ls), np.float64)
    lib().orc_moments_exact(np.ascontiguousarray(G, np.int64), len(G), ls, len(ls), m, v, s1, s2)
    return m, v, s1, s2


def moments_exact_fft(G, ls):
    """Exact (mean, var) via integer autocorrelation; None if the FFT rounding
    check fails."""
    ls = _u32(ls)
    m = np.zeros(len(ls), np.float64)
    v = np.zeros(len(ls), np.float64)
    ok = lib().orc_moments_exact_fft(np.ascontiguousarray(G, np.int64), len(G), ls, len(ls), m, v)
    return (m, v) if ok else None


def chrom_moments_exact(start, end, n_bins):
    """Exact compact table (mean, var) for one DB chromosome."""
    ls = compact_ls(n_bins)
    if len(ls) == 0:
        return ls, np.zeros(0), np.zeros(0)
    G = depth_exact(start, end, n_bins)
    if n_bins > 4096:
        r = moments_exact_fft(G, ls)
        if r is not None:
            return (ls,) + r
    m, v, _, _ = moments_exact(G, ls)
    return ls, m, v


# --------------------------------------------------------------------- stats

def erfc_as(x):
    return lib().orc_erfc_as(x)


def nb_llr(mu, var, obs):
    out = C.c_double()
    return out.value if lib().orc_nb_llr(mu, var, obs, C.byref(out)) else None


def poisson_llr(mu, obs):
    out = C.c_double()
    return out.value if lib().orc_poisson_llr(mu, obs, C.byref(out)) else None


def group_lengths(len_bp):
    len_bp = _u32(len_bp)
    n = len(len_bp)
    li = np.zeros(max(n, 1), np.uint32)
    s = np.zeros(max(n, 1), np.float64)
    s2 = np.zeros(max(n, 1), np.float64)
    c = np.zeros(max(n, 1), np.uint64)
    ng = lib().orc_group_lengths(n, len_bp, li, s, s2, c)
    return li[:ng], s[:ng], s2[:ng], c[:ng]


def analytic_stats(groups, slots, tables, observed):
    """compute_analytic_stats.  groups = (l_int, sum_scale, sum_scale2, count,
    chrom_slot); tables = list over slots of (n_bins, mean[], var[]).
    Returns None or (p, llr-or-None)."""
    li, s, s2, c, slot = groups
    nb = np.array([t[0] for t in tables] or [0], np.uint32)
    lens = np.array([len(t[1]) for t in tables] or [0], np.uint64)
    off = np.zeros(max(len(tables), 1), np.uint64)
    if len(tables):
        off[1:len(tables)] = np.cumsum(lens)[:-1]
    tm = np.concatenate([np.asarray(t[1], np.float64) for t in tables] + [np.zeros(1)])
    tv = np.concatenate([np.asarray(t[2], np.float64) for t in tables] + [np.zeros(1)])
    p, llr, has = C.c_double(), C.c_double(), C.c_int()
    ok = lib().orc_analytic_stats(len(li), _u32(li), np.ascontiguousarray(s, np.float64),
                                  np.ascontiguousarray(s2, np.float64),
                                  np.ascontiguousarray(c, np.uint64),
                                  np.ascontiguousarray(slot, np.int32), nb, off, lens, tm, tv,
                                  float(observed), C.byref(p), C.byref(llr), C.byref(has))
    if not ok:
        return None
    return p.value, (llr.value if has.value else None)
