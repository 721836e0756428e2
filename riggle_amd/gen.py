"""Seeded synthetic interval workloads of the benchmark configurations.

Genomes:
  * "c1": chr1 50 Mb, chr2 30 Mb, chr3 20 Mb.
  * "c2": 24 chromosomes, len_i = round(250e6 * 0.2**(i/23)), i = 0..23,
    named chr1..chr22, chrX, chrY (about 3.0 Gb).
Intervals: per-chromosome counts proportional to length; lengths
round(exp(ln(median) + sigma*Z)) clipped to [1, 100000]; a Bernoulli fraction
of intervals starts uniformly inside one of the genome's hotspot windows
(per chromosome k = max(1, round(1000 * len/G)) windows, together covering 5 %
of the chromosome, one placed uniformly inside each of k equal slots so they
never overlap); the rest start uniformly on [0, len - l).  Every draw comes
from numpy's counter-based Philox generator keyed by (seed, stream), so a
workload is reproducible from its seed.
"""
from __future__ import annotations

import numpy as np

CHROM_NAMES_24 = [f"chr{i}" for i in range(1, 23)] + ["chrX", "chrY"]


def genome(name: str):
    if name == "c1":
        return [("chr1", 50_000_000), ("chr2", 30_000_000), ("chr3", 20_000_000)]
    if name == "c2":
        return [(CHROM_NAMES_24[i], int(round(250e6 * 0.2 ** (i / 23)))) for i in range(24)]
    raise ValueError(name)


def _rng(seed: int, stream: int):
    return np.random.Generator(np.random.Philox(key=[seed & 0xFFFFFFFFFFFFFFFF, stream]))


def hotspots(genome_sizes, seed: int, cover=0.05):
    """Per chromosome: array of hotspot window starts and the window width."""
    G = sum(s for _, s in genome_sizes)
    rng = _rng(seed, 0xA5A5)
    out = []
    for _, size in genome_sizes:
        k = max(1, int(round(1000 * size / G)))
        width = max(1, int(cover * size / k))
        slot = size // k
        starts = np.arange(k, dtype=np.int64) * slot + rng.integers(0, max(1, slot - width), k)
        out.append((starts, width))
    return out


def _shard_of(cum, x):
    """np.searchsorted(cum, x, side="right") for x in [0, 1) through a 2^16-cell
    lookup table (exact whenever no cell holds two boundaries)."""
    M = 1 << 16
    if len(cum) > 1 and np.diff(cum).min() <= 2.0 / M:
        return np.searchsorted(cum, x, side="right")
    table = np.searchsorted(cum, np.arange(M, dtype=np.float64) / M, side="right")
    ext = np.append(cum, np.inf)
    t = table[(x * M).astype(np.int64)]
    return t + (ext[t] <= x)


def intervals(genome_sizes, n, seed, stream, median=1000.0, sigma=1.2, clip=(1, 100_000),
              hotspot_frac=0.0, hot=None):
    """n intervals: (shard index uint32, start uint32, end uint32)."""
    rng = _rng(seed, stream)
    sizes = np.array([s for _, s in genome_sizes], np.int64)
    p = sizes / sizes.sum()
    shard = _shard_of(np.cumsum(p), rng.random(n)).astype(np.uint32)
    shard = np.minimum(shard, len(sizes) - 1)
    ln = np.rint(np.exp(np.log(median) + sigma * rng.standard_normal(n)))
    ln = np.clip(ln, clip[0], clip[1]).astype(np.int64)
    csize = sizes[shard]
    ln = np.minimum(ln, csize - 1)
    start = (rng.random(n) * (csize - ln)).astype(np.int64)
    if hotspot_frac > 0 and hot is not None:
        is_hot = rng.random(n) < hotspot_frac
        idx = np.nonzero(is_hot)[0]
        hs = shard[idx]
        u = rng.random(len(idx))
        w = rng.random(len(idx))
        # every chromosome's windows in one array: interval i picks window
        # floor(u * k_c) of its chromosome and starts floor(w * width_c) into it
        nwin = np.array([len(ws) for ws, _ in hot], np.int64)
        woff = np.concatenate([[0], np.cumsum(nwin)[:-1]])
        width = np.array([wd for _, wd in hot], np.int64)
        allws = np.concatenate([ws for ws, _ in hot]).astype(np.int64)
        k = nwin[hs]
        wi = np.minimum((u * k.astype(np.float64)).astype(np.int64), k - 1)
        hstart = allws[woff[hs] + wi] + (w * width[hs].astype(np.float64)).astype(np.int64)
        hstart = np.minimum(hstart, csize[idx] - ln[idx])
        start[idx] = hstart
    return shard, start.astype(np.uint32), (start + ln).astype(np.uint32)


def log_uniform_counts(nfiles, lo, hi, seed, stream=0xC0C0):
    rng = _rng(seed, stream)
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), nfiles))).astype(np.int64)


def sort_files(shard, start, end, fid=None):
    """Each file's intervals in (chromosome, start) order, the files in fid
    order (as the generators lay them out)."""
    f = np.zeros(len(start), np.uint64) if fid is None else fid.astype(np.uint64)
    key = (f << np.uint64(37)) | (shard.astype(np.uint64) << np.uint64(32)) | start.astype(np.uint64)
    o = np.argsort(key)
    if fid is None:
        return shard[o], start[o], end[o]
    return shard[o], start[o], end[o], fid[o]


def make_db(genome_sizes, counts, seed, hotspot_frac, median, sigma, hot=None, sort=True):
    """Files with the given interval counts, each sorted by (chromosome,
    start) unless sort is False; returns (shard, start, end, fid)."""
    counts = np.asarray(counts, np.int64)
    N = int(counts.sum())
    shard, start, end = intervals(genome_sizes, N, seed, 1, median, sigma,
                                  hotspot_frac=hotspot_frac, hot=hot)
    fid = np.repeat(np.arange(len(counts), dtype=np.uint32), counts)
    return sort_files(shard, start, end, fid) if sort else (shard, start, end, fid)


def make_db_chunked(genome_sizes, counts, seed, hotspot_frac, median, sigma, hot=None,
                    chunk=1 << 25, threads=8):
    """make_db for billions of intervals: chunk c is drawn from stream
    (1 << 20) + c, chunks in parallel threads, into preallocated uint32
    arrays (shard, start, end, fid).  The files are left unsorted (sorting
    2e9 intervals on the host would dominate the run; the index build sorts
    them anyway)."""
    import concurrent.futures as cf
    counts = np.asarray(counts, np.int64)
    N = int(counts.sum())
    shard = np.empty(N, np.uint32)
    start = np.empty(N, np.uint32)
    end = np.empty(N, np.uint32)

    def work(c):
        a = c * chunk
        n = min(chunk, N - a)
        sh, s, e = intervals(genome_sizes, n, seed, (1 << 20) + c, median, sigma,
                             hotspot_frac=hotspot_frac, hot=hot)
        shard[a:a + n] = sh
        start[a:a + n] = s
        end[a:a + n] = e

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(work, range((N + chunk - 1) // chunk)))
    fid = np.repeat(np.arange(len(counts), dtype=np.uint32), counts)
    return shard, start, end, fid


def config(name: str, seed: int = 0xC40C1E000, sort: bool = True):
    """Returns a dict with genome, db arrays, query arrays for a benchmark config.
    sort=False leaves the C2 and C4 files in generation order (the same
    intervals; no result depends on the order, and sorting 1.45e8 intervals
    on the host costs about 15 s)."""
    if name == "c1":
        g = genome("c1")
        db = make_db(g, np.full(100, 10_000), seed + 1, 0.0, 500.0, 1.0)
        q = sort_files(*intervals(g, 10_000, seed + 1, 2, 500.0, 1.0))
        return {"genome": g, "db": db, "nfiles": 100, "query": q + (np.zeros(10_000, np.uint32),),
                "nq": 1}
    if name == "c2":
        g = genome("c2")
        hot = hotspots(g, seed + 2)
        counts = log_uniform_counts(1000, 1e3, 1e6, seed + 2)
        db = make_db(g, counts, seed + 2, 0.5, 1000.0, 1.2, hot, sort=sort)
        q = intervals(g, 1_000_000, seed + 2, 2, 1000.0, 1.2, hotspot_frac=0.5, hot=hot)
        q = sort_files(*q) if sort else q
        return {"genome": g, "db": db, "nfiles": 1000,
                "query": q + (np.zeros(1_000_000, np.uint32),), "nq": 1}
    if name == "c4":
        g = genome("c2")
        hot = hotspots(g, seed + 2)
        counts = log_uniform_counts(1000, 1e3, 1e6, seed + 2)
        db = make_db(g, counts, seed + 2, 0.5, 1000.0, 1.2, hot, sort=sort)
        qc = log_uniform_counts(256, 100, 2e6, seed + 4, stream=0xC4C4)
        sh, s, e = intervals(g, int(qc.sum()), seed + 4, 2, 1000.0, 1.2, hotspot_frac=0.3, hot=hot)
        qf = np.repeat(np.arange(256, dtype=np.uint32), qc)
        if sort:
            sh, s, e, qf = sort_files(sh, s, e, qf)
        return {"genome": g, "db": db, "nfiles": 1000, "query": (sh, s, e, qf), "nq": 256}
    if name == "c5":
        # serving: 10,000 files log-uniform 10k-1M (about 2.15e9 intervals),
        # 40 % hotspot; one sample query of 1M intervals (the serving leg
        # draws its own request stream)
        g = genome("c2")
        hot = hotspots(g, seed + 5)
        counts = log_uniform_counts(10_000, 1e4, 1e6, seed + 5)
        db = make_db_chunked(g, counts, seed + 5, 0.4, 1000.0, 1.2, hot)
        q = sort_files(*intervals(g, 1_000_000, seed + 5, 2, 1000.0, 1.2, hotspot_frac=0.4, hot=hot))
        return {"genome": g, "db": db, "nfiles": 10_000, "hot": hot,
                "query": q + (np.zeros(1_000_000, np.uint32),), "nq": 1}
    raise ValueError(name)
