"""The oracle against the reference's own golden vectors, known-answer tests
and closed forms (CPU only)."""
import numpy as np
import pytest

from tests.cases import flatten, integration_setup, random_files


def test_chacha20_published_vector(orc):
    # ChaCha20, all-zero key and nonce, block 0 (Bernstein's test vector TC1)
    b = orc.chacha_block(np.zeros(8, np.uint32), 0, 0, 20)
    assert [hex(x) for x in b[:4]] == ["0xade0b876", "0x903df1a0", "0xe56a5d40", "0x28bd8653"]
    assert hex(b[15]) == "0x8665eeb2"


def test_golden_intervals_hits_and_bp(orc):
    # tests/integration.rs:57-75: 158 hits, 413,652 bp
    db, q = integration_setup(orc, 10, 2000, 1, 1000, 50_000_000)
    c, bp = orc.sweep_bruteforce(db, q, 10, 1)
    assert c.sum() == 158 and bp.sum() == 413_652


def test_golden_counts_nonzero_pairs(orc):
    # tests/integration.rs:103-115: nnz 10, total 58
    db, q = integration_setup(orc, 5, 1000, 2, 500, 30_000_000)
    c, _ = orc.sweep_bruteforce(db, q, 5, 2)
    assert (c > 0).sum() == 10 and c.sum() == 58


def test_layered_reference_path_matches_exact(orc):
    db, q = integration_setup(orc, 5, 1000, 2, 500, 30_000_000)
    L = orc.LayeredDB(*db, 24, 5)
    cnt, bins, _ = L.query(*q, 2)
    c, bp = orc.sweep_bruteforce(db, q, 5, 2)
    assert (cnt == c).all()
    assert np.allclose(bins, bp / 100.0, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("seed", [0, 1])
def test_rank_identity_matches_bruteforce(orc, seed):
    rng = np.random.default_rng(seed)
    db = flatten(random_files(rng, 6, 3, (0, 300), 10_000, 1_500))
    q = flatten(random_files(rng, 3, 3, (0, 200), 10_000, 2_000))
    a = orc.sweep_bruteforce(db, q, 6, 3)
    b = orc.sweep_ranks(db, q, 6, 3, 3)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


def test_layer_config(orc):
    # src/io/meta.rs:152-170
    assert [orc.layer_for_size(s) for s in (0, 127, 128, 511, 512, 2047, 2048)] == [0, 0, 1, 1, 2, 2, 3]
    assert [orc.layer_max_size(k) for k in (0, 1, 2)] == [128, 512, 2048]


def test_jump_table(orc):
    # src/io/jump.rs:181-240
    assert orc.build_jump_table([], 64).tolist() == [0]
    assert orc.build_jump_table([10, 50, 80], 100).tolist() == [0]
    assert orc.build_jump_table([10, 150, 250], 100).tolist() == [0, 1, 2]
    assert orc.build_jump_table([10, 350], 100).tolist() == [0, 1, 1, 1]


def test_stats_known_answers(orc):
    # src/fourier/stats.rs:218-242
    assert abs(orc.nb_llr(4000.0, 8000.0, 4000.0)) < 1e-6
    assert orc.nb_llr(4000.0, 8000.0, 20_000.0) > 0
    assert orc.nb_llr(2000.0, 1000.0, 1.0) is None
    assert abs(orc.erfc_as(0.0) - 1.0) < 1e-6
    assert abs(orc.erfc_as(1.0) - 0.1572992) < 1e-5
    assert orc.erfc_as(5.0) < 1e-10
    # A&S 7.1.26 against the true erfc (error < 1.5e-7)
    import math
    for x in (0.1, 0.5, 1.0, 2.0, 3.0):
        assert abs(orc.erfc_as(x) - math.erfc(x)) < 1.5e-7


def test_stats_fixed_moments(orc):
    # stats.rs:197-216: one 10-bin interval, moments (20, 100) everywhere
    g = orc.group_lengths([1000])
    groups = g + (np.zeros(1, np.int32),)
    table = [(508_185, np.full(2000, 20.0), np.full(2000, 100.0))]
    p0, _ = orc.analytic_stats(groups, None, table, 0.0)
    assert p0 == 1.0
    p10, llr = orc.analytic_stats(groups, None, table, 10.0)
    assert 0.5 < p10 < 1.0 and llr is not None


def test_compact_table_layout(orc):
    ls = orc.compact_ls(1_000_000)
    assert list(ls[:100]) == list(range(1, 101))
    # every stored L maps to its own slot; intermediate L to the slot below
    for i, L in enumerate(ls):
        assert orc.compact_index(int(L)) == i
    assert orc.compact_index(102) == 100 and orc.compact_index(103) == 101


def test_depth_and_moment_closed_forms(orc):
    # depth.rs:222-230: [20,80) covers 0.6 of a bin
    assert orc.depth_signal_f32([20], [80], 10)[0] == pytest.approx(0.6)
    assert orc.depth_exact([20], [80], 10)[0] == 60
    # moments.rs:474-488 autocorrelation known answers
    assert orc.autocorr_fft(np.array([1.0, 2, 3, 0, 0]))[0] == pytest.approx(14.0)
    assert orc.autocorr_fft(np.array([0.0, 1, 1, 1, 2, 2, 2, 3]))[1] == pytest.approx(18.0)
    # moments.rs:511-562: uniform g = 1 -> mean L, var 0; periodic [1,0,0]
    G = np.full(50, 100, np.int64)
    m, v, _, _ = orc.moments_exact(G, np.arange(1, 50))
    assert np.allclose(m, np.arange(1, 50)) and np.all(v == 0)
    m, v, _, _ = orc.moments_exact(np.array([100, 0, 0, 100, 0, 0, 100]), [3, 1])
    assert m[0] == pytest.approx(1.0) and v[0] == 0 and m[1] == pytest.approx(3 / 7)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_exact_moments_against_window_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    nb = 3000
    s = rng.integers(0, nb * 100, 400).astype(np.uint32)
    e = (s + rng.integers(1, 20_000, 400)).astype(np.uint32)
    G = orc.depth_exact(s, e, nb)
    ls = orc.compact_ls(nb)
    m1, v1, _, _ = orc.moments_exact(G, ls)
    m2, v2 = orc.moments_exact_fft(G, ls)
    assert np.array_equal(m1, m2) and np.array_equal(v1, v2)
    P = np.concatenate([[0], np.cumsum(G)])
    for i in (0, 7, 99, len(ls) - 1):
        L = int(ls[i])
        w = (P[L:] - P[:-L]).astype(np.float64) / 100.0
        assert m1[i] == pytest.approx(w.mean(), rel=1e-12)
        assert v1[i] == pytest.approx(w.var(), rel=1e-9, abs=1e-9)


def test_stdrng_range_i64_power_of_two_closed_form(orc):
    # for range = 2^k the zone is (2^k << (63 - k)) - 1 = 2^63 - 1: a draw v
    # is kept when bit 63 - k of v is clear (the low word of v * 2^k is below
    # 2^63) and gives lo + (v >> (64 - k)); pins the widening multiply and the
    # rejection rule
    r1, r2 = orc.StdRng(99), orc.StdRng(99)
    rejected = 0
    for k in (1, 7, 31, 40) * 8:
        lo = -(1 << (k - 1))
        got = r1.range_i64(lo, lo + (1 << k) - 1)
        while True:
            v = r2.next_u64()
            if not (v >> (63 - k)) & 1:
                break
            rejected += 1
        assert got == lo + (v >> (64 - k))
    assert rejected > 0


def test_stdrng_range_i64_bounds_and_spread(orc):
    r = orc.StdRng(5)
    m = 248_956_422
    v = [r.range_i64(-m, m) for _ in range(2000)]
    assert min(v) >= -m and max(v) <= m
    assert min(v) < -m // 2 and max(v) > m // 2 and abs(sum(v) / len(v)) < m / 10


@pytest.mark.parametrize("seed", [3, 4])
def test_layered_exact_scan_matches_bruteforce(orc, seed):
    # the reference's windowed scan with exact accumulators equals the pair
    # brute force (and the rank identity) on random multi-file cases
    from tests.cases import flatten, random_files
    rng = np.random.default_rng(seed)
    S, F, nq = 3, 9, 70
    db = flatten(random_files(rng, F, S, (20, 300), 400_000, 40_000))
    q = flatten(random_files(rng, nq, S, (0, 40), 400_000, 60_000))
    L = orc.LayeredDB(*db, S, F)
    c, b = L.query_exact(*q, nq, nthreads=3)
    oc, ob = orc.sweep_bruteforce(db, q, F, nq)
    rc, rb = orc.sweep_ranks(db, q, F, nq, S)
    assert (c == oc).all() and (b == ob).all()
    assert (rc == oc).all() and (rb == ob).all()
