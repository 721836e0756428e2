"""Precision form of the moments: the FP32 lag-limited FFT route
(chk_moments_widths) against the exact-integer oracle at the configurations'
16 widths (1..1000 bins) and at edge cases.

The oracle's moments_exact evaluates the reference method
(src/fourier/moments.rs:192-254) with exact integer window sums; the
tolerance is north_star's 1e-6 relative on every mean and variance (1e-12
absolute floor for zero variances).  The lags themselves are checked against
exact integer autocorrelations on short tracks."""
import numpy as np
import pytest

from riggle_amd import gen
from tests.cases import flatten, random_files

pytestmark = pytest.mark.gpu

C3_WIDTHS = [int(round(10 ** (k / 5))) for k in range(16)]
REL = 1e-6


def check(orc, db, F, S, sizes, widths, m, v):
    sh, s, e, fid = db
    worst_m = worst_v = 0.0
    checked = 0
    for f in range(F):
        for c in range(S):
            nb = (int(sizes[c]) + 99) // 100
            sel = (fid == f) & (sh == c)
            if not sel.any() or nb < 2:
                assert np.isnan(m[f, c]).all() and np.isnan(v[f, c]).all(), (f, c)
                continue
            G = orc.depth_exact(s[sel], e[sel], nb)
            ok = [i for i, L in enumerate(widths) if L < nb]
            bad = [i for i, L in enumerate(widths) if L >= nb]
            assert np.isnan(m[f, c, bad]).all()
            xm, xv, _, _ = orc.moments_exact(G, np.array([widths[i] for i in ok], np.uint32))
            gm, gv = m[f, c, ok], v[f, c, ok]
            worst_m = max(worst_m, float(np.max(np.abs(gm - xm) / np.maximum(np.abs(xm), 1e-12))))
            worst_v = max(worst_v, float(np.max(np.abs(gv - xv) / np.maximum(np.abs(xv), 1e-12))))
            checked += 1
    return worst_m, worst_v, checked


def test_c3_law_tracks_at_the_16_widths(chk, orc):
    g = gen.genome("c2")
    S = len(g)
    sizes = np.array([x for _, x in g], np.uint64)
    seed = 0xC40C1E000 + 3
    db = gen.make_db(g, [1_000, 30_000, 1_000_000], seed, 0.5, 1000.0, 1.2, gen.hotspots(g, seed))
    ix = chk.Index(*db, S, 3)
    ix.set_genome(sizes)
    m, v, info = ix.moments_widths(C3_WIDTHS)
    assert info["inexact_rows"] == 0
    print("exact (signal, width) pairs:", info["exact_widths"], "of", 3 * S * 16)
    wm, wv, n = check(orc, db, 3, S, sizes, C3_WIDTHS, m, v)
    assert n == 3 * S
    assert wm <= REL and wv <= REL, (wm, wv)


@pytest.mark.parametrize("seed", [0, 1])
def test_edges_short_tracks_and_lags(chk, orc, seed):
    # shard lengths around the 3072-bin hop and the 4096-point block, tracks
    # shorter than the widths, deep coverage, intervals past the table end
    rng = np.random.default_rng(seed)
    bins = [2, 3, 57, 1024, 1025, 3071, 3072, 3073, 6144, 6145, 20_000, 99_999]
    sizes = np.array([100 * b - int(rng.integers(0, 99)) for b in bins], np.uint64)
    S, F = len(bins), 4
    files = []
    for f in range(F):
        sh, st, en = [], [], []
        for c, size in enumerate(sizes):
            k = int(rng.integers(1, 200)) * (f + 1)
            a = rng.integers(0, int(size) + 500, k)
            ln = rng.integers(1, 30_000 if f < 3 else 300, k)
            sh.append(np.full(k, c))
            st.append(a)
            en.append(a + ln)
        files.append((np.concatenate(sh), np.concatenate(st), np.concatenate(en)))
    db = flatten(files)
    ix = chk.Index(*db, S, F)
    ix.set_genome(sizes)
    widths = [1, 2, 3, 5, 100, 999, 1000, 1024]
    m, v, info, lags = ix.moments_widths(widths, lags=True)
    wm, wv, n = check(orc, db, F, S, sizes, widths, m, v)
    assert n == F * S and wm <= REL and wv <= REL, (wm, wv)
    # lags against the exact integer autocorrelation of the centred coverage
    sh, s, e, fid = db
    for f, c in ((0, 3), (1, 7), (3, 9)):
        nb = bins[c]
        sel = (fid == f) & (sh == c)
        G = orc.depth_exact(s[sel], e[sel], nb)
        H = G - int(np.floor(G.sum() / nb + 0.5))  # llround of the mean, as the device centres
        ex = np.array([int(np.dot(H[:nb - d], H[d:])) if d < nb else 0 for d in range(1024)], np.float64)
        scale = max(1.0, float(ex[0]))
        assert np.max(np.abs(lags[f, c] - ex)) <= 1e-6 * scale, (f, c)


def test_widths_out_of_range_are_rejected(chk):
    z = np.zeros(4, np.uint32)
    ix = chk.Index(z, np.arange(4, dtype=np.uint32) * 100, np.arange(4, dtype=np.uint32) * 100 + 50,
                   z, 1, 1)
    ix.set_genome(np.array([10_000], np.uint64))
    with pytest.raises(chk.ChuckleError, match="widths"):
        ix.moments_widths([1, 1025])
    with pytest.raises(chk.ChuckleError, match="widths"):
        ix.moments_widths([0])
