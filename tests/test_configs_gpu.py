"""GPU parity on the benchmark workloads drawn by riggle_amd.gen (the
configurations' genomes, file-size laws, hotspots and length laws):

  C1  the whole workload: 1 query file of 10,000 intervals vs 100 files x
      10,000 (counts and bp vs the oracle's rank sweep), the moment tables of
      every chromosome of 8 files vs the exact-integer oracle, p-values of all
      100 cells vs the oracle's FP64 statistics on the device's tables;
  C2  the whole workload: 1,000,000 query intervals vs the 1,000-file DB of
      about 1.45e8 intervals, counts and bp bit-exact vs the rank sweep;
  C3  C2-law tracks on the four shortest C2 chromosomes: moment tables vs
      the exact oracle at every stored block size, the 16 C3 widths included;
  C4  256 query files (100-20,000 intervals, 30 % hotspot) vs the C2 DB
      through one captured HIP graph, vs the reference's windowed scan with
      exact accumulators;
  C5  a serving mix drawn with C5's laws (files log-uniform, 40 % hotspot,
      Poisson arrivals, sizes log-uniform from 1,000) against a C5-law DB of
      120 files, every row vs the rank sweep.

Tolerances: counts and bp bit-exact; moments 1e-6 relative; p-values 1e-3
in -log10 p and LLR 1e-9 relative (FP64 summation order only)."""
import threading
import time

import numpy as np
import pytest

from riggle_amd import gen
from tests.stats_cases import mlog10, oracle_cell_stats

pytestmark = pytest.mark.gpu

C3_WIDTHS = [int(round(10 ** (k / 5))) for k in range(16)]


def sizes_of(genome):
    return np.array([s for _, s in genome], np.uint64)


def tables_from_gpu(ix, F, S, sizes):
    """The device's compact tables as the oracle's per-(file, shard) tables."""
    so, sl, m, v = ix.download_moments()
    tabs = {}
    for f in range(F):
        for c in range(S):
            k = f * S + c
            if so[k] >= 0:
                tabs[(f, c)] = ((int(sizes[c]) + 99) // 100, m[so[k]:so[k] + sl[k]],
                                v[so[k]:so[k] + sl[k]])
    return tabs


def check_tables(orc, db, tabs, f, c, nb):
    sh, s, e, fid = db
    sel = (fid == f) & (sh == c)
    ls, xm, xv = orc.chrom_moments_exact(s[sel], e[sel], nb)
    _, gm, gv = tabs[(f, c)]
    assert len(gm) == len(ls)
    np.testing.assert_allclose(gm, xm, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(gv, xv, rtol=1e-6, atol=1e-12)
    return ls


def test_c1_whole_workload(chk, orc):
    cfg = gen.config("c1")
    db, q = cfg["db"], cfg["query"]
    S, F = len(cfg["genome"]), cfg["nfiles"]
    sizes = sizes_of(cfg["genome"])
    ix = chk.Index(*db, S, F)
    ix.set_genome(sizes)
    rows, exact = ix.build_moments()
    assert rows == F * S
    out = ix.stats(ix.query(q[0], q[1], q[2]))
    oc, obp = orc.sweep_ranks(db, q, F, 1, S)
    assert (out["counts"].astype(np.int64) == oc).all()
    assert (out["bp"].astype(np.int64) == obp).all()
    tabs = tables_from_gpu(ix, F, S, sizes)
    for f in range(0, F, 14):  # every chromosome of 8 files (the oracle's FFT is slow)
        for c in range(S):
            check_tables(orc, db, tabs, f, c, (int(sizes[c]) + 99) // 100)
    emitted = 0
    for f in range(F):
        exp = oracle_cell_stats(orc, q, 0, f, S, sizes, tabs, obp[0, f] / 100.0) \
            if oc[0, f] > 0 else None
        assert bool(out["flags"][0, f] & 1) == (exp is not None), f
        if exp is None:
            continue
        emitted += 1
        p, llr = exp
        assert bool(out["flags"][0, f] & 2) == (llr is not None)
        assert abs(mlog10(out["p"][0, f]) - mlog10(p)) < 1e-3
        if llr is not None:
            assert out["llr"][0, f] == pytest.approx(llr, rel=1e-9, abs=1e-9)
    assert emitted == F


@pytest.fixture(scope="module")
def c2(chk):
    cfg = gen.config("c2", sort=False)
    ix = chk.Index(*cfg["db"], len(cfg["genome"]), cfg["nfiles"])
    return cfg, ix


def test_c2_whole_workload(chk, orc, c2):
    cfg, ix = c2
    S, F = len(cfg["genome"]), cfg["nfiles"]
    q = cfg["query"]
    assert len(cfg["db"][1]) > 1.0e8 and len(q[1]) == 1_000_000
    cnt, bp = ix.sweep(ix.query(q[0], q[1], q[2]))
    oc, obp = orc.sweep_ranks(cfg["db"], q, F, 1, S)
    assert (cnt.astype(np.int64) == oc).all()
    assert (bp.astype(np.int64) == obp).all()
    assert int(cnt.sum()) > 1e9  # the hotspots make it a pair-heavy query


def test_c4_reduced_batch_graph(chk, orc, c2):
    cfg, ix = c2
    S, F = len(cfg["genome"]), cfg["nfiles"]
    g = cfg["genome"]
    hot = gen.hotspots(g, 0xC40C1E000 + 2)
    qc = gen.log_uniform_counts(256, 100, 20_000, 0xC40C1E000 + 4, stream=0xC4C4)
    sh, s, e = gen.intervals(g, int(qc.sum()), 0xC40C1E000 + 4, 2, 1000.0, 1.2,
                             hotspot_frac=0.3, hot=hot)
    qf = np.repeat(np.arange(256, dtype=np.uint32), qc)
    qq = ix.query(sh, s, e, qf, 256)
    info = qq.info()
    dc, db_ = ix.alloc(8 * 256 * F), ix.alloc(8 * 256 * F)
    graph = ix.graph(qq, dc, db_)
    for _ in range(2):
        graph.launch()
        ix.sync()
    cnt = dc.to_numpy(np.uint64, (256, F))
    bp = db_.to_numpy(np.uint64, (256, F))
    L = orc.LayeredDB(*cfg["db"], S, F, nthreads=16)
    oc, obp = L.query_exact(sh, s, e, qf, 256, nthreads=16)
    assert (cnt.astype(np.int64) == oc).all()
    assert (bp.astype(np.int64) == obp).all()
    assert info["nfiles"] == 256 and info["dense_files"] < 256


def test_c3_tracks_at_c3_widths(chk, orc):
    g = gen.genome("c2")
    S = len(g)
    sizes = sizes_of(g)
    seed = 0xC40C1E000 + 3
    hot = gen.hotspots(g, seed)
    counts = [1_000, 30_000, 1_000_000]
    db = gen.make_db(g, counts, seed, 0.5, 1000.0, 1.2, hot)
    ix = chk.Index(*db, S, len(counts))
    ix.set_genome(sizes)
    rows, exact = ix.build_moments()
    tabs = tables_from_gpu(ix, len(counts), S, sizes)
    checked = 0
    for f in range(len(counts)):
        for c in range(S - 4, S):  # chr21, chr22, chrX and chrY: the shortest
            if (f, c) not in tabs:
                continue
            ls = check_tables(orc, db, tabs, f, c, (int(sizes[c]) + 99) // 100)
            assert set(C3_WIDTHS[:11]) <= set(ls.tolist())  # L <= 100 are stored exactly
            checked += 1
    assert checked >= 9


def test_c5_law_serving_mix(chk, orc):
    g = gen.genome("c2")
    S = len(g)
    seed = 0xC40C1E000 + 5
    hot = gen.hotspots(g, seed)
    F = 120
    counts = gen.log_uniform_counts(F, 1e4, 1e5, seed)
    db = gen.make_db(g, counts, seed, 0.4, 1000.0, 1.2, hot)
    ix = chk.Index(*db, S, F)
    rng = np.random.default_rng(5)
    nreq = 40
    sizes = gen.log_uniform_counts(nreq, 1e3, 2e5, seed, stream=0xC5C5)
    reqs = [gen.intervals(g, int(n), seed, 0x5000 + i, 1000.0, 1.2, hotspot_frac=0.4, hot=hot)
            for i, n in enumerate(sizes)]
    srv = chk.Server(ix, slots=3, max_batch_files=8)
    tickets = []

    def submitter():
        for r in reqs:  # Poisson arrivals, mean gap 2 ms
            time.sleep(rng.exponential(0.002))
            tickets.append(srv.submit(*r))

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    rows = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    qsh = np.concatenate([r[0] for r in reqs])
    qs = np.concatenate([r[1] for r in reqs])
    qe = np.concatenate([r[2] for r in reqs])
    qf = np.repeat(np.arange(nreq, dtype=np.uint32), [len(r[1]) for r in reqs])
    oc, obp = orc.sweep_ranks(db, (qsh, qs, qe, qf), F, nreq, S)
    for i, r in enumerate(rows):
        assert (r["counts"].astype(np.int64) == oc[i]).all(), i
        assert (r["bp"].astype(np.int64) == obp[i]).all(), i
    assert st["completed"] == nreq and st["p99_ms"] >= st["p50_ms"] > 0
