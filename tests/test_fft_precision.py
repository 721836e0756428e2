"""Why the moment FFT runs in FP64: on a synthetic chromosome of the C2 shape
(coverage of 100-bp bins from the C2 interval law, hotspots included), an FP64
FFT autocorrelation lies within 0.25 of the exact integers and rounds to them,
so every moment is exact up to the final division.  An FP32 transform's
errors, carried into the variance of the widest windows (the compact tables
store block sizes up to n - 1 bins), exceed the 1e-6 relative tolerance of
the moment tables.  CPU only (scipy.fft in FP64 and FP32)."""
import numpy as np
import scipy.fft as sf

from riggle_amd import gen


def _signal():
    g = gen.genome("c2")
    hot = gen.hotspots(g, 0xC40C1E000 + 2)
    sh, s, e = gen.intervals(g, 600_000, 0xC40C1E000 + 2, 7, 1000.0, 1.2, hotspot_frac=0.5, hot=hot)
    k = 5  # a 176 Mb chromosome: 1.76 M bins
    s, e = s[sh == k].astype(np.int64), e[sh == k].astype(np.int64)
    n = (g[k][1] + 99) // 100
    d = np.zeros(n * 100 + 1, np.int64)
    np.add.at(d, s, 1)
    np.add.at(d, e, -1)
    G = np.cumsum(d)[: n * 100].reshape(n, 100).sum(1)
    return G - int(round(G.sum() / n))


def _autocorr(h, dtype):
    n = len(h)
    N = 1 << int(np.ceil(np.log2(2 * n - 1)))
    X = sf.rfft(h.astype(dtype), N)
    return sf.irfft(X * np.conj(X), N)[:n].astype(np.float64)


def _var_rel_err(h, dA, Ls):
    # the variance of window sums takes dA through S2 = L A0 + 2 sum_{k<L} (L - k) A[k] + ...
    n = len(h)
    ch = np.concatenate([[0], np.cumsum(h)])
    out = []
    for L in Ls:
        nw = n - L + 1
        S = (ch[L:] - ch[:-L]).astype(np.longdouble)
        var = float(((S * S).sum() / nw - (S.sum() / nw) ** 2) / 1e4)
        k = np.arange(L, dtype=np.float64)
        dT = L * dA[0] + 2.0 * np.sum((L - k[1:]) * dA[1:L])
        out.append(abs(dT / (1e4 * nw) / var))
    return np.array(out)


def test_fp64_rounds_exactly_and_fp32_misses_the_tolerance():
    h = _signal()
    n = len(h)
    exact = np.rint(_autocorr(h, np.float64))
    a64 = _autocorr(h, np.float64)
    assert exact[0] == float((h.astype(object) ** 2).sum())  # A[0] = sum H^2
    assert np.abs(a64 - exact).max() < 0.25                  # rounds to the exact integers
    Ls = [1, 100, 10_000, n // 2, n - n // 20, n - 100]
    rel32 = _var_rel_err(h, _autocorr(h, np.float32) - exact, Ls)
    assert rel32[:3].max() < 1e-6   # narrow windows would pass in FP32 ...
    assert rel32.max() > 1e-6       # ... the widest stored ones would not
