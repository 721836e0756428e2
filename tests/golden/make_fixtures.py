"""Writes the golden fixtures of tests/golden/ from first principles, with no
use of the oracle or of the reference's code:

  sweep_cases.json   overlap counts / bp by brute force over all pairs
                     (half-open rule d.start < q.end && d.end > q.start)
  moment_cases.json  sliding-window coverage moments from direct window sums
                     of the integer bp coverage per 100-bp bin, exact
                     rational arithmetic (fractions.Fraction)
  stats_cases.json   NB / Poisson log-likelihood ratios from scipy's
                     nbinom / poisson log-pmf differences (the LLR is the
                     log-pmf at the fitted parameter minus the null's), and
                     right tails 0.5 * erfc(sqrt(LLR)) with the C library erfc

Run: python tests/golden/make_fixtures.py  (seeded; rewrites the files)
"""
import json
import math
import os
import random
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def brute(db, q, F, nq):
    c = [[0] * F for _ in range(nq)]
    b = [[0] * F for _ in range(nq)]
    for (qs_, qa, qe, r) in q:
        for (ds_, da, de, f) in db:
            if qs_ == ds_ and da < qe and de > qa:
                c[r][f] += 1
                b[r][f] += min(qe, de) - max(qa, da)
    return c, b


def sweep_cases():
    rng = random.Random(20250101)
    cases = []

    def rand_iv(n, S, span, maxlen, nfiles):
        out = []
        for _ in range(n):
            s = rng.randrange(S)
            a = rng.randrange(span)
            out.append([s, a, a + rng.randint(1, maxlen), rng.randrange(nfiles)])
        return out

    for (S, span, maxlen, F, nq, ndb, nqv) in [(1, 1000, 50, 3, 1, 60, 20), (2, 5000, 400, 7, 3, 200, 60),
                                              (3, 20000, 3000, 12, 2, 300, 90),
                                              (1, 200, 200, 5, 4, 150, 40)]:
        db = rand_iv(ndb, S, span, maxlen, F)
        q = rand_iv(nqv, S, span, maxlen, nq)
        c, b = brute(db, q, F, nq)
        cases.append({"nshards": S, "F": F, "nq": nq, "db": db, "q": q, "counts": c, "bp": b})
    # hand-made edges: touching, nesting, identical, coordinate 0 and 2^31-1
    top = 2**31 - 1
    db = [[0, 0, 10, 0], [0, 10, 20, 1], [0, 5, 15, 2], [0, 5, 15, 2], [0, top - 5, top, 3],
          [0, 0, top, 4], [0, 100, 101, 1]]
    q = [[0, 10, 11, 0], [0, 0, 5, 0], [0, 5, 15, 1], [0, top - 1, top, 1], [0, 101, 200, 0],
         [0, 99, 100, 1]]
    c, b = brute(db, q, 5, 2)
    cases.append({"nshards": 1, "F": 5, "nq": 2, "db": db, "q": q, "counts": c, "bp": b})
    return cases


def coverage(ivs, n):
    G = [0] * n
    lim = 100 * n
    for a, e in ivs:
        a, e = min(a, lim), min(e, lim)
        for j in range(a // 100, (e + 99) // 100):
            lo, hi = max(a, 100 * j), min(e, 100 * j + 100)
            if hi > lo:
                G[j] += hi - lo
    return G


def window_moments(G, L):
    n = len(G)
    P = [0]
    for x in G:
        P.append(P[-1] + x)
    nw = n - L + 1
    sums = [P[k + L] - P[k] for k in range(nw)]
    S1 = sum(sums)
    S2 = sum(s * s for s in sums)
    mean = Fraction(S1, 100 * nw)
    var = Fraction(S2, 10_000 * nw) - mean * mean
    return float(mean), float(max(var, Fraction(0)))


def moment_cases():
    rng = random.Random(77)
    cases = []

    def track(n_bins, nint, maxlen):
        span = 100 * n_bins
        ivs = []
        for _ in range(nint):
            a = rng.randrange(span + 300)  # some intervals reach past the end
            ivs.append([a, a + rng.randint(1, maxlen)])
        return ivs

    specs = [("random_short", 300, 80, 400), ("random_long", 1500, 500, 3000),
             ("deep", 200, 400, 2000), ("sub_bin", 120, 60, 40)]
    for name, n, k, ml in specs:
        cases.append((name, n, track(n, k, ml)))
    cases.append(("full", 150, [[0, 15_000]]))
    cases.append(("empty", 50, []))
    cases.append(("two_bins", 2, [[30, 170]]))
    out = []
    for name, n, ivs in cases:
        G = coverage(ivs, n)
        Ls = sorted(set([1, 2, 3, 5, 10, 17, 50, 99, 100, n // 3, n // 2, n - 2, n - 1]))
        Ls = [L for L in Ls if 1 <= L <= n - 1]
        ms = [window_moments(G, L) for L in Ls]
        out.append({"name": name, "n_bins": n, "intervals": ivs, "G": G, "L": Ls,
                    "mean": [m for m, _ in ms], "var": [v for _, v in ms]})
    return out


def stats_cases():
    from scipy.stats import nbinom, poisson
    rows = []
    for mu, var, obs in [(40.0, 90.0, 80), (40.0, 90.0, 10), (500.0, 4000.0, 2500), (7.5, 30.0, 1),
                         (1000.0, 1500.0, 1300), (3.0, 12.0, 40)]:
        r = mu * mu / (var - mu)
        p = mu / var
        p_o = r / (r + obs)
        llr = float(nbinom.logpmf(obs, r, p_o) - nbinom.logpmf(obs, r, p))
        rows.append({"kind": "nb", "mu": mu, "var": var, "obs": obs, "llr": llr,
                     "tail": 0.5 * math.erfc(math.sqrt(max(llr, 0.0)))})
    for mu, obs in [(20.0, 35), (20.0, 8), (300.0, 420), (2.5, 9)]:
        llr = float(poisson.logpmf(obs, obs) - poisson.logpmf(obs, mu))
        rows.append({"kind": "poisson", "mu": mu, "obs": obs, "llr": llr,
                     "tail": 0.5 * math.erfc(math.sqrt(max(llr, 0.0)))})
    return rows


if __name__ == "__main__":
    for name, data in (("sweep_cases.json", sweep_cases()), ("moment_cases.json", moment_cases()),
                       ("stats_cases.json", stats_cases())):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print("wrote", name)
