"""GPU compact moment tables vs the exact-integer oracle.

The oracle evaluates the reference method (src/fourier/moments.rs:192-254) in
exact integers; the tolerance is 1e-6 relative on every mean and variance
(1e-12 absolute floor for zero variances).  Shard lengths cover both FFT
forms: one LDS-resident FFT (n <= 4096 bins) and the four-step passes."""
import numpy as np
import pytest

from tests.cases import flatten, random_files
from tests.stats_cases import oracle_tables

pytestmark = pytest.mark.gpu

REL = 1e-6


def build(chk, db, F, S, sizes):
    ix = chk.Index(*db, S, F)
    ix.set_genome(sizes)
    rows, exact = ix.build_moments()
    return ix, rows, exact


def compare(ix, orc, db, F, S, sizes):
    so, sl, m, v = ix.download_moments()
    oso, osl, om, ov, tabs = oracle_tables(orc, db, F, S, sizes)
    assert np.array_equal(sl, osl)
    assert np.array_equal(so >= 0, oso >= 0)
    worst_m = worst_v = 0.0
    for k in np.nonzero(so >= 0)[0]:
        a, b, L = so[k], oso[k], sl[k]
        gm, gv, xm, xv = m[a:a + L], v[a:a + L], om[b:b + L], ov[b:b + L]
        if L == 0:
            continue
        worst_m = max(worst_m, float(np.max(np.abs(gm - xm) / np.maximum(np.abs(xm), 1e-12))))
        worst_v = max(worst_v, float(np.max(np.abs(gv - xv) / np.maximum(np.abs(xv), 1e-12))))
    return worst_m, worst_v


@pytest.mark.parametrize("seed,span", [(0, 300_000), (1, 3_000_000), (2, 20_000_000)])
def test_tables_match_exact_oracle(chk, orc, seed, span):
    rng = np.random.default_rng(seed)
    F, S = 5, 3
    sizes = np.array([span, span // 2 + 77, 0], np.uint64)
    db = flatten(random_files(rng, F, S, (1, 3000), span, 20_000))
    ix, rows, exact = build(chk, db, F, S, sizes)
    assert rows > 0
    wm, wv = compare(ix, orc, db, F, S, sizes)
    assert wm <= REL and wv <= REL, (wm, wv)


def test_deep_coverage_and_edges(chk, orc):
    # depth > 1, intervals past the table end (clipped), sub-bin intervals
    rng = np.random.default_rng(9)
    S, F = 2, 2
    sizes = np.array([1_000_050, 5_000], np.uint64)
    s = rng.integers(0, 1_050_000, 4000).astype(np.uint32)
    e = (s + rng.integers(1, 200_000, 4000)).astype(np.uint32)
    sh = (rng.random(4000) < 0.9).astype(np.uint32) ^ 1
    db = (sh, s, e, rng.integers(0, F, 4000).astype(np.uint32))
    ix, rows, exact = build(chk, db, F, S, sizes)
    wm, wv = compare(ix, orc, db, F, S, sizes)
    assert wm <= REL and wv <= REL, (wm, wv)


def test_gpu_moments_drive_pvalues(chk, orc):
    from tests.stats_cases import oracle_cell_stats, mlog10
    rng = np.random.default_rng(4)
    F, nq, S = 4, 2, 3
    sizes = np.array([2_000_000, 1_000_000, 600_000], np.uint64)
    db = flatten(random_files(rng, F, S, (100, 600), 600_000, 20_000))
    q = flatten(random_files(rng, nq, S, (50, 300), 600_000, 20_000))
    ix, rows, exact = build(chk, db, F, S, sizes)
    out = ix.stats(ix.query(q[0], q[1], q[2], q[3], nq))
    _, _, _, _, tabs = oracle_tables(orc, db, F, S, sizes)
    oc, obp = orc.sweep_ranks(db, q, F, nq, S)
    n = 0
    for r in range(nq):
        for f in range(F):
            exp = oracle_cell_stats(orc, q, r, f, S, sizes, tabs, obp[r, f] / 100.0) \
                if oc[r, f] > 0 else None
            assert bool(out["flags"][r, f] & 1) == (exp is not None)
            if exp is not None:
                n += 1
                assert abs(mlog10(out["p"][r, f]) - mlog10(exp[0])) < 1e-3
    assert n > 3


def test_fft_pair_groups_match_single_launch(chk, orc):
    # four-step shard (n = 30,000 bins, N = 65,536) with 7 files: a 1 MiB FFT
    # workspace runs the 4 signal pairs one per launch, the last pair with a
    # missing odd row; tables equal the one-launch build bit for bit
    rng = np.random.default_rng(21)
    F, S = 7, 2
    sizes = np.array([3_000_000, 2_000], np.uint64)
    db = flatten(random_files(rng, F, S, (50, 3000), 3_000_000, 20_000))
    ix1, rows1, _ = build(chk, db, F, S, sizes)
    ix2 = chk.Index(*db, S, F)
    ix2.set_genome(sizes)
    ix2.tune("moments.fft_mib", 1)
    assert ix2.tune_get("moments.fft_mib") == 1
    rows2, _ = ix2.build_moments()
    assert rows1 == rows2 > 0
    a, b = ix1.download_moments(), ix2.download_moments()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    wm, wv = compare(ix2, orc, db, F, S, sizes)
    assert wm <= REL and wv <= REL, (wm, wv)


@pytest.mark.parametrize("block", [4096, 8192])
def test_fft_block_sizes_match_oracle(chk, orc, block):
    # shard 0: n = 30,000 bins (four-step either way, N1 = block / 2);
    # shard 1: n = 3,000 bins, N = 8192 (one block at 8192, four-step at 4096)
    rng = np.random.default_rng(33)
    F, S = 4, 2
    sizes = np.array([3_000_000, 300_000], np.uint64)
    db = flatten(random_files(rng, F, S, (50, 3000), 3_000_000, 20_000))
    ix = chk.Index(*db, S, F)
    ix.set_genome(sizes)
    ix.tune("moments.fft_block", block)
    rows, _ = ix.build_moments()
    assert rows > 0
    wm, wv = compare(ix, orc, db, F, S, sizes)
    assert wm <= REL and wv <= REL, (wm, wv)


def test_rounding_modes_and_rescan(chk, orc):
    # mode 1 never rounds; mode 2 starts every row rounded and, since FFT
    # outputs are not exact integers, rescans it unrounded: its tables equal
    # mode 1's bit for bit; all three match the oracle
    rng = np.random.default_rng(41)
    F, S = 3, 1
    sizes = np.array([3_000_000], np.uint64)
    db = flatten(random_files(rng, F, S, (400, 3000), 3_000_000, 50_000))
    out = {}
    for mode in (0, 1, 2):
        ix = chk.Index(*db, S, F)
        ix.set_genome(sizes)
        ix.tune("moments.round", mode)
        rows, exact = ix.build_moments()
        out[mode] = (rows, exact, ix.download_moments())
        wm, wv = compare(ix, orc, db, F, S, sizes)
        assert wm <= REL and wv <= REL, (mode, wm, wv)
    assert out[0][1] == out[0][0] == F
    assert out[1][1] == 0 and out[2][1] == 0
    for x, y in zip(out[1][2], out[2][2]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", [0, 2])
def test_eval_segments_match_oracle(chk, orc, mode):
    # shard 0: n = 200,000 bins, evaluated in 4 segments by default and in 49
    # at 4,096 bins; mode 2 scans every row's A side a second time, unrounded,
    # over all its segments.  Both segmentations match the oracle and each
    # other (means bit for bit: they are exact integers over the window)
    rng = np.random.default_rng(57)
    F, S = 3, 2
    sizes = np.array([20_000_000, 1_000_000], np.uint64)
    db = flatten(random_files(rng, F, S, (200, 3000), 20_000_000, 30_000))
    tabs = []
    for seg in (0, 4096):
        ix = chk.Index(*db, S, F)
        ix.set_genome(sizes)
        ix.tune("moments.round", mode)
        ix.tune("moments.eval_seg", seg)
        assert ix.tune_get("moments.eval_seg") == seg
        rows, exact = ix.build_moments()
        assert rows > 0 and exact == (rows if mode == 0 else 0)
        wm, wv = compare(ix, orc, db, F, S, sizes)
        assert wm <= REL and wv <= REL, (seg, wm, wv)
        tabs.append(ix.download_moments())
    (so0, sl0, m0, v0), (so1, sl1, m1, v1) = tabs
    assert np.array_equal(so0, so1) and np.array_equal(sl0, sl1)
    assert np.array_equal(m0, m1)
    np.testing.assert_allclose(v1, v0, rtol=1e-12, atol=1e-12)


def test_deep_rounded_rows_match_oracle(chk, orc):
    # 750 intervals stacked over the first half of a 3 Mb shard: bins of
    # 75,000 bp coverage against a few hundred, sum H^2 ~ 4.2e13 above 2^45
    # and sum G ~ 1.1e9 above 2^29, so that row's scans sum their runs in
    # double-double (the shallow row takes the FP64 and int64 run sums);
    # both rows stay exact
    rng = np.random.default_rng(61)
    F, S = 2, 1
    sizes = np.array([3_000_000], np.uint64)
    files = random_files(rng, F, S, (100, 500), 3_000_000, 20_000)
    deep = (np.zeros(750, np.uint32), np.zeros(750, np.uint32), np.full(750, 1_500_000, np.uint32))
    files[0] = tuple(np.concatenate([x, y]) for x, y in zip(files[0], deep))
    db = flatten(files)
    ix, rows, exact = build(chk, db, F, S, sizes)
    assert rows == exact == F
    wm, wv = compare(ix, orc, db, F, S, sizes)
    assert wm <= REL and wv <= REL, (wm, wv)


def test_segment_and_chunk_edges_match_oracle(chk, orc):
    # chromosomes of 2, 3, 4,095, 4,096, 4,097 and 8,192 bins, evaluated in
    # 4,096-position segments and in the default ones: positions 0..n give a
    # segment holding only position n (n = 4,096 and 8,192), exactly one
    # full chunk (n = 4,095), a last chunk of two positions (n = 4,097) and
    # the two-bin minimum table (L = 1 only)
    rng = np.random.default_rng(71)
    nb = [2, 3, 4095, 4096, 4097, 8192]
    sizes = np.array([100 * k - 50 for k in nb], np.uint64)
    S, F = len(nb), 3
    files = []
    for _ in range(F):
        sh, s, e = [], [], []
        for c, size in enumerate(int(x) for x in sizes):
            k = int(rng.integers(1, 200))
            st = rng.integers(0, size, k)
            ln = rng.integers(1, max(2, size // 3), k)
            sh.append(np.full(k, c, np.uint32))
            s.append(st.astype(np.uint32))
            e.append(np.minimum(st + ln, size).astype(np.uint32))
        files.append((np.concatenate(sh), np.concatenate(s), np.concatenate(e)))
    db = flatten(files)
    for seg in (4096, 0):
        ix = chk.Index(*db, S, F)
        ix.set_genome(sizes)
        ix.tune("moments.eval_seg", seg)
        rows, exact = ix.build_moments()
        assert rows == exact == S * F
        wm, wv = compare(ix, orc, db, F, S, sizes)
        assert wm <= REL and wv <= REL, (seg, wm, wv)


def test_fully_covered_chromosome_mean(chk, orc):
    # src/fourier/moments.rs:580-592: one interval covering chr22 (50,818,468 bp)
    # gives mean(L) ~ L bins of coverage; the reference checks lookup(500)
    # within 2 %.  The entry compact_index(500) holds the stored block size
    # Ls next to it on the x1.01 ladder, and only the window over the partial
    # last bin (68 bp) falls short, so mean(Ls) = Ls within 1e-8
    size = 50_818_468
    one = np.zeros(1, np.uint32)
    db = (one, one, np.array([size], np.uint32), one)
    ix, rows, exact = build(chk, db, 1, 1, [size])
    so, sl, m, v = ix.download_moments()
    k = orc.compact_index(500)
    ls = orc.compact_ls((size + 99) // 100)
    assert so[0] >= 0 and k < sl[0] == len(ls)
    mean = m[so[0] + k]
    assert abs(mean - 500.0) / 500.0 < 0.02
    assert mean == pytest.approx(float(ls[k]), rel=1e-8)
    assert 0.0 <= v[so[0] + k] < 1e-4
    worst_m, worst_v = compare(ix, orc, db, 1, 1, [size])
    assert worst_m <= REL and worst_v <= REL


def test_forward_column_pass_forms_agree(chk, orc):
    # the persistent forward column pass (default) and the one-block-per-group
    # kernel (moments.fft_pipe = 1) give the same tables bit for bit.  Shards
    # of 5,000, 30,000, 100,000 and 300,000 bins: column FFTs of 8 points
    # (always the per-group kernel), 32 (radix-4 first pass), 128 (radix-2) and
    # 512 (radix-8; 384 column groups, so persistent blocks take several);
    # 5 files, so the last signal pair has no odd row
    rng = np.random.default_rng(45)
    F, S = 5, 4
    sizes = np.array([500_000, 3_000_000, 10_000_000, 30_000_000], np.uint64)
    files = []
    for _ in range(F):
        n = int(rng.integers(2000, 6000))
        sh = rng.integers(0, S, n).astype(np.uint32)
        st = (rng.random(n) * (sizes[sh].astype(np.float64) - 30_000)).astype(np.uint32)
        files.append((sh, st, st + rng.integers(1, 30_000, n).astype(np.uint32)))
    db = flatten(files)
    tabs = []
    for pipe in (0, 1):
        ix = chk.Index(*db, S, F)
        ix.set_genome(sizes)
        ix.tune("moments.fft_pipe", pipe)
        assert ix.tune_get("moments.fft_pipe") == pipe
        rows, _ = ix.build_moments()
        assert rows > 0
        tabs.append(ix.download_moments())
        if pipe == 0:
            wm, wv = compare(ix, orc, db, F, S, sizes)
            assert wm <= REL and wv <= REL, (wm, wv)
    for x, y in zip(*tabs):
        assert np.array_equal(x, y)
