"""Parity of the HIP merge-path sweep with the oracle (bit-exact u64).

Mirrors the reference's sweep tests (src/sweep/query.rs:159-221), its seeded
integration goldens (tests/integration.rs:57-115) and adds randomized cases
against the oracle's brute-force and rank-identity forms.
"""
import numpy as np
import pytest

from tests.cases import flatten, integration_setup, random_files

pytestmark = pytest.mark.gpu


def gpu_sweep(chk, db, q, nshards, F, nq):
    """Sweep with the dense merge-path path and with the sparse pair path;
    both must agree exactly before either is compared with the oracle."""
    ix = chk.Index(db[0], db[1], db[2], db[3], nshards, F)
    out = []
    for method in (1, 2):
        ix.tune("sweep.method", method)
        qq = ix.query(q[0], q[1], q[2], q[3] if nq > 1 else None, nq)
        out.append(ix.sweep(qq))
    (c1, b1), (c2, b2) = out
    assert (c1 == c2).all() and (b1 == b2).all(), "dense and sparse paths disagree"
    return c1, b1


def one(shard_start_end_sid):
    return tuple(np.array(v, np.uint32) for v in zip(*shard_start_end_sid))


def run_unit(chk, db, q, nq, F):
    db = one(db) if db else tuple(np.zeros(0, np.uint32) for _ in range(4))
    q = one(q) if q else tuple(np.zeros(0, np.uint32) for _ in range(4))
    return gpu_sweep(chk, db, q, 1, F, nq)


# --- src/sweep/query.rs:158-221 ------------------------------------------------

def test_basic_overlap(chk):
    c, bp = run_unit(chk, [(0, 50, 150, 0)], [(0, 40, 160, 0)], 1, 1)
    assert c[0, 0] == 1 and bp[0, 0] == 100  # 1.0 bin


def test_no_overlap(chk):
    c, bp = run_unit(chk, [(0, 50, 100, 0)], [(0, 200, 300, 0)], 1, 1)
    assert c[0, 0] == 0 and bp[0, 0] == 0


def test_multiple_queries_and_db(chk):
    c, bp = run_unit(chk, [(0, 25, 75, 0), (0, 75, 125, 1)], [(0, 0, 50, 0), (0, 50, 100, 1)], 2, 2)
    assert c.tolist() == [[1, 0], [1, 1]]
    assert bp.tolist() == [[25, 0], [25, 25]]


def test_dead_zone(chk):
    c, bp = run_unit(chk, [(0, 0, 50, 0), (0, 10_000, 10_050, 1)], [(0, 10_000, 10_100, 0)], 1, 2)
    assert c.tolist() == [[0, 1]] and bp.tolist() == [[0, 50]]


def test_touching_intervals_not_overlapping(chk):
    c, bp = run_unit(chk, [(0, 0, 50, 0)], [(0, 50, 100, 0)], 1, 1)
    assert c[0, 0] == 0 and bp[0, 0] == 0
    c, bp = run_unit(chk, [(0, 50, 100, 0)], [(0, 0, 50, 0)], 1, 1)
    assert c[0, 0] == 0 and bp[0, 0] == 0


def test_empty_db(chk):
    c, bp = run_unit(chk, [], [(0, 0, 100, 0)], 1, 1)
    assert c[0, 0] == 0 and bp[0, 0] == 0


def test_empty_query(chk):
    c, bp = run_unit(chk, [(0, 0, 100, 0)], [], 1, 1)
    assert c[0, 0] == 0 and bp[0, 0] == 0


def test_shard_not_in_db(chk):
    # src/tasks/query/mod.rs:393-412: a query shard absent from the DB gives nothing
    db = one([(0, 100, 200, 0)])
    q = one([(7, 100, 200, 0)])  # id >= nshards: shard unknown to the DB
    c, bp = gpu_sweep(chk, db, q, 1, 1, 1)
    assert c[0, 0] == 0 and bp[0, 0] == 0


# --- tests/integration.rs goldens ---------------------------------------------

def test_golden_total_hits_and_bp(chk, orc):
    db, q = integration_setup(orc, 10, 2000, 1, 1000, 50_000_000)
    c, bp = gpu_sweep(chk, db, q, 24, 10, 1)
    assert int(c.sum()) == 158 and int(bp.sum()) == 413_652


def test_golden_counts_nonzero_pairs(chk, orc):
    db, q = integration_setup(orc, 5, 1000, 2, 500, 30_000_000)
    c, bp = gpu_sweep(chk, db, q, 24, 5, 2)
    assert int((c > 0).sum()) == 10 and int(c.sum()) == 58
    oc, obp = orc.sweep_bruteforce(db, q, 5, 2)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


# --- randomized parity ----------------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_vs_bruteforce(chk, orc, seed):
    rng = np.random.default_rng(seed)
    F, nq, S = 7, 3, 4
    db = flatten(random_files(rng, F, S, (0, 400), 20_000, 2_000))
    q = flatten(random_files(rng, nq, S + 1, (0, 300), 20_000, 3_000))  # shard S unknown to DB
    c, bp = gpu_sweep(chk, db, q, S, F, nq)
    oc, obp = orc.sweep_bruteforce(db, q, F, nq)
    assert (c.astype(np.int64) == oc).all()
    assert (bp.astype(np.int64) == obp).all()


def test_ties_and_nesting(chk, orc):
    # identical coordinates, zero-gap touching, nested, 1-bp intervals
    rng = np.random.default_rng(9)
    n = 3000
    s = rng.integers(0, 60, n).astype(np.uint32) * 10
    e = s + rng.integers(1, 8, n).astype(np.uint32) * 10
    e[::7] = s[::7] + 1
    db = (np.zeros(n, np.uint32), s, e, rng.integers(0, 5, n).astype(np.uint32))
    qs = rng.integers(0, 60, 500).astype(np.uint32) * 10
    q = (np.zeros(500, np.uint32), qs, qs + rng.integers(1, 6, 500).astype(np.uint32) * 10,
         rng.integers(0, 2, 500).astype(np.uint32))
    c, bp = gpu_sweep(chk, db, q, 1, 5, 2)
    oc, obp = orc.sweep_bruteforce(db, q, 5, 2)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


def test_large_vs_ranks(chk, orc):
    rng = np.random.default_rng(11)
    F, nq, S = 40, 4, 6
    db = flatten(random_files(rng, F, S, (1000, 12_000), 5_000_000, 20_000))
    q = flatten(random_files(rng, nq, S, (10, 30_000), 5_000_000, 20_000))
    c, bp = gpu_sweep(chk, db, q, S, F, nq)
    oc, obp = orc.sweep_ranks(db, q, F, nq, S)
    assert (c.astype(np.int64) == oc).all()
    assert (bp.astype(np.int64) == obp).all()


def test_query_dense_tiles(chk, orc):
    # a query much larger than the DB: most tiles hold only query events
    rng = np.random.default_rng(5)
    db = flatten(random_files(rng, 3, 2, 50, 1_000_000, 5_000))
    q = flatten(random_files(rng, 1, 2, 200_000, 1_000_000, 300))
    c, bp = gpu_sweep(chk, db, q, 2, 3, 1)
    oc, obp = orc.sweep_ranks(db, q, 3, 1, 2)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


def test_repeat_is_deterministic(chk, orc):
    rng = np.random.default_rng(3)
    db = flatten(random_files(rng, 20, 3, 5000, 2_000_000, 5_000))
    q = flatten(random_files(rng, 2, 3, 4000, 2_000_000, 5_000))
    ix = chk.Index(*db, 3, 20)
    qq = ix.query(q[0], q[1], q[2], q[3], 2)
    a = ix.sweep(qq)
    b = ix.sweep(qq)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


@pytest.mark.parametrize("chunk,items,bpc", [(1, 4, 1), (3, 8, 2), (64, 16, 0), (2, 8, 0), (5, 12, 0)])
def test_launch_shapes_agree(chk, orc, chunk, items, bpc):
    # small chunks and tiles force work-queue items that span several jobs,
    # query-file switches inside one block and the pipeline's odd tail
    rng = np.random.default_rng(21)
    F, nq, S = 12, 5, 7
    db = flatten(random_files(rng, F, S, (2000, 9000), 3_000_000, 8_000))
    q = flatten(random_files(rng, nq, S, (100, 20_000), 3_000_000, 8_000))
    ix = chk.Index(*db, S, F)
    ix.tune("sweep.method", 1)
    ix.tune("sweep.chunk", chunk)
    ix.tune("sweep.items", items)
    ix.tune("sweep.blocks_per_cu", bpc)
    c, bp = ix.sweep(ix.query(q[0], q[1], q[2], q[3], nq))
    oc, obp = orc.sweep_ranks(db, q, F, nq, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


@pytest.mark.parametrize("spi", [1, 7, 256])
def test_sparse_item_sizes_agree(chk, orc, spi):
    rng = np.random.default_rng(22)
    F, nq, S = 9, 4, 3
    db = flatten(random_files(rng, F, S, (1000, 5000), 1_000_000, 50_000))
    q = flatten(random_files(rng, nq, S, (1, 900), 1_000_000, 50_000))
    ix = chk.Index(*db, S, F)
    ix.tune("sweep.method", 2)
    ix.tune("sweep.sparse_items", spi)
    c, bp = ix.sweep(ix.query(q[0], q[1], q[2], q[3], nq))
    oc, obp = orc.sweep_ranks(db, q, F, nq, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


def test_mixed_batch_chooses_per_file(chk, orc):
    # file sizes straddle the sparse/dense threshold: both kernels write
    # rows of the same output in one sweep
    rng = np.random.default_rng(23)
    F, S = 30, 4
    db = flatten(random_files(rng, F, S, (500, 4000), 4_000_000, 20_000))
    sizes = [3, 40, 400, 4000, 20_000, 1, 90_000]
    q = flatten([random_files(rng, 1, S, n, 4_000_000, 20_000)[0] for n in sizes])
    ix = chk.Index(*db, S, F)
    ix.tune("sweep.sparse_max", 500)
    qq = ix.query(q[0], q[1], q[2], q[3], len(sizes))
    info = qq.info()  # files above 500 intervals go dense
    assert info["nfiles"] == len(sizes) and info["n"] == sum(sizes)
    assert info["dense_files"] == 3 and info["dense_intervals"] == 4000 + 20_000 + 90_000
    c, bp = ix.sweep(qq)
    oc, obp = orc.sweep_ranks(db, q, F, len(sizes), S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


@pytest.mark.parametrize("F", [1500, 2600])
def test_many_files_use_file_groups(chk, orc, F):
    # more DB files than one LDS histogram holds: the dense streams and the
    # layered table are split into groups of 1024 files, each with its own
    # LDS histogram
    rng = np.random.default_rng(24 + F)
    S = 2
    db = flatten(random_files(rng, F, S, (5, 60), 2_000_000, 30_000))
    q = flatten(random_files(rng, 3, S, (2000, 6000), 2_000_000, 30_000))
    c, bp = gpu_sweep(chk, db, q, S, F, 3)
    oc, obp = orc.sweep_ranks(db, q, F, 3, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


def test_lengths_across_all_layers(chk, orc):
    # DB lengths from 1 bp to 5 Mb populate many size layers; the sparse
    # path's look-back per layer must still find every overlap
    rng = np.random.default_rng(25)
    F, S, span = 6, 2, 50_000_000
    files = []
    for _ in range(F):
        n = 3000
        ln = np.exp(rng.uniform(0, np.log(5e6), n)).astype(np.uint32) + 1
        s = rng.integers(0, span, n).astype(np.uint32)
        files.append((rng.integers(0, S, n).astype(np.uint32), s, s + ln))
    db = flatten(files)
    q = flatten(random_files(rng, 2, S, 2000, span, 200_000))
    c, bp = gpu_sweep(chk, db, q, S, F, 2)
    oc, obp = orc.sweep_ranks(db, q, F, 2, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


def test_tune_rejects_unknown_key(chk):
    ix = chk.Index(np.zeros(1, np.uint32), np.zeros(1, np.uint32), np.ones(1, np.uint32),
                   np.zeros(1, np.uint32), 1, 1)
    with pytest.raises(chk.ChuckleError):
        ix.tune("sweep.nonsense", 1)


def _shift(t, off):
    return t[0], t[1] + np.uint32(off), t[2] + np.uint32(off), t[3]


def test_coordinates_near_int32_max(chk, orc):
    # event keys 2x + b then need all 32 bits: the widest coordinate field of
    # the event sort; one interval ends exactly at 2^31 - 1
    rng = np.random.default_rng(31)
    F, nq, S, L, span = 6, 3, 3, 5_000, 200_000
    off = 2**31 - 1 - span - L
    db = _shift(flatten(random_files(rng, F, S, (500, 2000), span, L)), off)
    q = _shift(flatten(random_files(rng, nq, S, (200, 1000), span, L)), off)
    q = tuple(np.append(a, v).astype(np.uint32) for a, v in zip(q, (0, 2**31 - 50, 2**31 - 1, 0)))
    assert int(q[2].max()) == 2**31 - 1
    c, bp = gpu_sweep(chk, db, q, S, F, nq)
    oc, obp = orc.sweep_bruteforce(db, q, F, nq)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


def test_many_segments_sorted_in_five_passes(chk, orc):
    # 48 dense query files x 30 shards = 1,440 event segments (11 bits) above
    # 32-bit coordinates: 43 key bits, sorted in passes of 9, 9, 9, 8 and 8
    rng = np.random.default_rng(32)
    F, nq, S, L, span = 5, 48, 30, 3_000, 1_000_000
    off = 2**31 - 1 - span - L
    db = _shift(flatten(random_files(rng, F, S, (3000, 6000), span, L)), off)
    q = _shift(flatten(random_files(rng, nq, S, (300, 800), span, L)), off)
    ix = chk.Index(*db, S, F)
    ix.tune("sweep.method", 1)
    c, bp = ix.sweep(ix.query(q[0], q[1], q[2], q[3], nq))
    oc, obp = orc.sweep_ranks(db, q, F, nq, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()


def test_single_interval_query(chk, orc):
    # two events: one tile of the event sort, most of it past the end
    db = (np.array([0, 0], np.uint32), np.array([10, 500], np.uint32),
          np.array([600, 900], np.uint32), np.array([0, 1], np.uint32))
    q = tuple(np.array([v], np.uint32) for v in (0, 550, 700, 0))
    c, bp = gpu_sweep(chk, db, q, 1, 2, 1)
    assert c.tolist() == [[1, 1]] and bp.tolist() == [[50, 150]]


def test_query_reset_reuses_handle(chk, orc):
    # one handle, batches of changing size and file count: every reset gives
    # the oracle's matrix
    rng = np.random.default_rng(91)
    S, F = 3, 40
    db = flatten(random_files(rng, F, S, (100, 3000), 2_000_000, 20_000))
    ix = chk.Index(*db, S, F)
    q = None
    for nq, per in ((3, (500, 4000)), (1, 90_000), (7, (10, 300)), (2, (20_000, 60_000))):
        qq = flatten(random_files(rng, nq, S, per, 2_000_000, 30_000))
        if q is None:
            q = ix.query(qq[0], qq[1], qq[2], qq[3], nq)
        else:
            q.reset(qq[0], qq[1], qq[2], qq[3], nq)
        c, bp = ix.sweep(q)
        oc, obp = orc.sweep_ranks(db, qq, F, nq, S)
        assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all(), nq


@pytest.mark.parametrize("gap", [3_000_000, 5_000_000, 70_000_000])
def test_sparse_and_clustered_streams(chk, orc, gap):
    # DB streams of very different densities, across two file groups: shard 0
    # spaces its intervals `gap` apart up to the top of the int32 range,
    # shard 1 mixes a dense 1 Mb cluster with far outliers, shard 2 stays
    # inside 2 Mb; dense method at two tile sizes against the rank sweep
    rng = np.random.default_rng(gap % 1_000_003)
    F, S = 1100, 3
    n0 = min(400, (2**31 - 200_000) // gap)
    s0 = (np.arange(n0, dtype=np.int64) * gap + rng.integers(0, 1000, n0)).astype(np.uint32)
    cl = rng.integers(50_000_000, 51_000_000, 6000).astype(np.uint32)
    far = rng.integers(0, 2**31 - 200_000, 300).astype(np.uint32)
    s1 = np.concatenate([cl, far])
    s2 = rng.integers(0, 2_000_000, 5000).astype(np.uint32)
    st = np.concatenate([s0, s1, s2])
    ln = rng.integers(1, 100_000, len(st)).astype(np.uint32)
    sh = np.concatenate([np.zeros(n0), np.ones(len(s1)), np.full(len(s2), 2)]).astype(np.uint32)
    fid = rng.integers(0, F, len(st)).astype(np.uint32)
    db = (sh, st, st + ln, fid)
    nq = 3
    qs = np.concatenate([rng.integers(0, 2**31 - 2_100_000, 3000), rng.integers(49_000_000, 52_000_000, 3000),
                         rng.integers(0, 2_100_000, 3000)]).astype(np.uint32)
    qsh = np.repeat(np.arange(3), 3000).astype(np.uint32)
    q = (qsh, qs, qs + rng.integers(1, 2_000_000, len(qs)).astype(np.uint32),
         rng.integers(0, nq, len(qs)).astype(np.uint32))
    for items in (4, 8):
        ix = chk.Index(*db, S, F)
        ix.tune("sweep.method", 1)
        ix.tune("sweep.items", items)
        c, bp = ix.sweep(ix.query(q[0], q[1], q[2], q[3], nq))
        oc, obp = orc.sweep_ranks(db, q, F, nq, S)
        assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all(), items


def test_large_batch_staged_over_host_threads(chk, orc):
    # 700k query intervals in 5 files (ids interleaved) with some on a shard the
    # DB lacks: staged over several host threads, the batch gives the oracle's
    # matrix; the first invalid interval in input order is the one reported
    rng = np.random.default_rng(77)
    F, S, nq, n = 30, 4, 5, 700_000
    db = flatten(random_files(rng, F, 3, (2_000, 20_000), 3_000_000, 5_000))
    st = rng.integers(0, 3_000_000, n).astype(np.uint32)
    q = (rng.integers(0, S, n).astype(np.uint32), st, st + rng.integers(1, 5_000, n).astype(np.uint32),
         rng.integers(0, nq, n).astype(np.uint32))
    ix = chk.Index(*db, S, F)
    c, bp = ix.sweep(ix.query(q[0], q[1], q[2], q[3], nq))
    keep = q[0] < 3
    oc, obp = orc.sweep_ranks(db, tuple(a[keep] for a in q), F, nq, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()
    bad = [a.copy() for a in q]
    bad[3][500_000] = nq          # file id out of range, later in the input
    bad[2][300_000] = bad[1][300_000]  # start == end, earlier
    with pytest.raises(chk.ChuckleError, match="start must be less than end"):
        ix.query(bad[0], bad[1], bad[2], bad[3], nq)
    bad[2][300_000] = bad[1][300_000] + 1
    with pytest.raises(chk.ChuckleError, match="file id out of range"):
        ix.query(bad[0], bad[1], bad[2], bad[3], nq)


def test_auto_plan_with_empty_and_sampled_files(chk, orc):
    # the per-file visit estimate picks each file's method: files of at most
    # 8,192 intervals are estimated whole, larger ones from a strided sample;
    # empty files (no intervals, or all on a shard the DB lacks) sit between
    # them, and two DB file groups double the sub-tables per interval
    rng = np.random.default_rng(29)
    F, S = 1100, 4
    db = flatten(random_files(rng, F, S, (20, 200), 2_000_000, 5_000))
    sizes = [0, 9000, 0, 20_000, 5, 8192, 8193, 1, 0]
    files = [random_files(rng, 1, S, n, 2_000_000, 5_000)[0] for n in sizes]
    files[0] = (np.full(300, S + 3, np.uint32), np.arange(300, dtype=np.uint32),
                np.arange(300, dtype=np.uint32) + 7)  # every interval on an absent shard
    q = flatten(files)
    ix = chk.Index(*db, S, F)
    qq = ix.query(q[0], q[1], q[2], q[3], len(sizes))
    info = qq.info()
    assert info["nfiles"] == len(sizes) and info["n"] == sum(sizes)
    c, bp = ix.sweep(qq)
    keep = q[0] < S
    oc, obp = orc.sweep_ranks(db, tuple(a[keep] for a in q), F, len(sizes), S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()
    assert not c[[0, 2, 8]].any() and not bp[[0, 2, 8]].any()


def _pileup_case(rng):
    # 40,000 query intervals on one shard, 24,000 of them starting at one
    # coordinate and 8,000 ending at another: long runs of equal event keys
    n, S, F = 40_000, 2, 9
    db = flatten(random_files(rng, F, S, (2000, 6000), 3_000_000, 30_000))
    s = rng.integers(0, 2_900_000, n).astype(np.uint32)
    s[:24_000] = 1_234_567
    e = s + rng.integers(1, 40_000, n).astype(np.uint32)
    e[30_000:38_000] = 2_950_000
    s[30_000:38_000] = np.minimum(s[30_000:38_000], 2_949_000)
    q = (np.zeros(n, np.uint32), s, e, np.zeros(n, np.uint32))
    return db, q, S, F, 1


def _sparse_segments_case(rng):
    # 40 query files x 25 shards with a few intervals each over 2^31
    # coordinates: 1,000 event segments above 32-bit coordinate keys, each
    # sort tile spanning many of them
    S, F, nq = 25, 6, 40
    db = flatten(random_files(rng, F, S, (300, 900), 2**31 - 100_000, 50_000))
    q = flatten(random_files(rng, nq, S, (20, 120), 2**31 - 100_000, 50_000))
    return db, q, S, F, nq


def _hotspot_case(rng):
    # three dense files with half of their intervals in 1 % of the span
    S, F, nq = 4, 12, 3
    db = flatten(random_files(rng, F, S, (5000, 20_000), 20_000_000, 5_000))
    files = []
    for _ in range(nq):
        n = 60_000
        sh = rng.integers(0, S, n).astype(np.uint32)
        st = rng.integers(0, 20_000_000, n).astype(np.uint32)
        hot = rng.random(n) < 0.5
        st[hot] = 9_000_000 + rng.integers(0, 200_000, int(hot.sum())).astype(np.uint32)
        files.append((sh, st, st + rng.integers(1, 3_000, n).astype(np.uint32)))
    return db, flatten(files), S, F, nq


@pytest.mark.parametrize("case", [_pileup_case, _sparse_segments_case, _hotspot_case],
                         ids=["pileup", "sparse_segments", "hotspot"])
def test_event_sort_skewed_inputs(chk, orc, case):
    # event distributions that stress the event sort: the dense sweep's
    # matrices match the oracle
    db, q, S, F, nq = case(np.random.default_rng(77))
    ix = chk.Index(*db, S, F)
    ix.tune("sweep.method", 1)
    c, bp = ix.sweep(ix.query(q[0], q[1], q[2], q[3] if nq > 1 else None, nq))
    oc, obp = orc.sweep_ranks(db, q, F, nq, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()
