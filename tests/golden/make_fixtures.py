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
  group_cases.json   query length groups (l_int = max(1, ceil(len / 100)),
                     sum of l / l_int, sum of (l / l_int)^2, count) in exact
                     rationals, with lengths on both sides of every ceil
                     boundary
  analytic_cases.json  whole cells of the analytic statistics: query
                     intervals per chromosome, per-chromosome moment tables
                     (block sizes 1..K, so table slot = L - 1), the observed
                     overlap in bins and the expected outcome: no record,
                     NB or Poisson branch, LLR from scipy log-pmf
                     differences at integer points, p from the C library
                     erfc on the enrichment or depletion side

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


def groups_exact(lengths_bp):
    """{l_int: [sum l/l_int, sum (l/l_int)^2, count]} with l = len/100 exact."""
    out = {}
    for b in lengths_bp:
        l = Fraction(b, 100)
        li = max(1, math.ceil(l))
        g = out.setdefault(li, [Fraction(0), Fraction(0), 0])
        g[0] += l / li
        g[1] += (l / li) ** 2
        g[2] += 1
    return out


def group_cases():
    specs = [[1, 99, 100, 101, 150, 200, 201],
             [1000],
             [50, 50, 100, 100, 100, 250, 300, 301, 99_999],
             [7, 7, 7, 7, 7, 12_345, 12_399, 12_400, 12_401],
             [201, 1, 100, 99, 150, 101, 200]]  # unsorted input
    rows = []
    for lens in specs:
        g = groups_exact(lens)
        rows.append({"len_bp": lens, "l_int": sorted(g),
                     "sum_scale": [float(g[k][0]) for k in sorted(g)],
                     "sum_scale2": [float(g[k][1]) for k in sorted(g)],
                     "count": [g[k][2] for k in sorted(g)]})
    return rows


def analytic_cases():
    """Cells of compute_analytic_stats (src/fourier/stats.rs:34-88) derived
    from the documented method: mu = sum mean(l_int) * sum_scale, var = sum
    var(l_int) * sum_scale2, mu_raw = sum mean(l_int) / l_int * count over
    groups whose table has l_int (l_int < n_bins, slot inside the table);
    no group -> no record; var > mu -> NB(r = mu^2 / (var - mu), p = mu / var)
    tilted to r / (r + O); else Poisson(mu_raw) at O / (mu / mu_raw)."""
    from scipy.stats import nbinom, poisson
    chroms = {"chrA": 100_000, "chrB": 50_000, "chrC": 500, "chrD": 100_000}

    def table(kind, K):
        Ls = range(1, K + 1)
        if kind == "const":
            return [2.0] * K, [1.0] * K
        if kind == "over":
            return [0.5 * L for L in Ls], [0.3 * L * L + L for L in Ls]
        if kind == "over2":
            return [0.2 * L for L in Ls], [2.0 * L for L in Ls]
        raise ValueError(kind)

    specs = [
        # the Poisson branch: 4 x 1000 bp (l_int 10, sum_scale 4, sum_scale2 4)
        # under mean 2, var 1: mu 8, var 4, mu_raw 0.8, effective scale 10
        ("poisson_enriched", {"chrA": [1000] * 4}, {"chrA": ("const", 100)}, 40),
        ("poisson_30", {"chrA": [1000] * 4}, {"chrA": ("const", 100)}, 30),
        # 20 x 100 bp (l_int 1): mu 40, var 20, mu_raw 40, effective scale 1
        ("poisson_depleted", {"chrA": [100] * 20}, {"chrA": ("const", 100)}, 15),
        ("nb_two_chroms", {"chrA": [150, 250, 201, 99, 1000, 100], "chrB": [100, 300]},
         {"chrA": ("over", 100), "chrB": ("over2", 100)}, 22),
        ("nb_depleted", {"chrA": [150, 250, 201, 99, 1000, 100], "chrB": [100, 300]},
         {"chrA": ("over", 100), "chrB": ("over2", 100)}, 3),
        # chrC has 5 bins: l_int 6 >= n_bins is skipped, l_int 2 is kept
        ("skip_lint_ge_nbins", {"chrC": [600, 200], "chrA": [300]},
         {"chrC": ("over", 4), "chrA": ("over", 100)}, 11),
        # a table of 8 slots has no entry for l_int 10: only l_int 3 counts
        ("skip_past_table", {"chrA": [1000, 300, 300]}, {"chrA": ("over", 8)}, 16),
        # the query's chromosome has no table for this DB file
        ("no_shared_chrom", {"chrD": [500, 700]}, {"chrA": ("over", 100)}, 12),
        # O = 0 under the NB branch: p_o = 1 fails the guard -> no LLR, p = 1
        ("nb_zero_observed", {"chrA": [150, 250, 201, 99, 1000, 100]}, {"chrA": ("over", 100)}, 0),
    ]
    rows = []
    for name, q, tabs, O in specs:
        mu = var = mu_raw = Fraction(0)
        shared = False
        for c, lens in q.items():
            if c not in tabs:
                continue
            n_bins = (chroms[c] + 99) // 100
            mean_t, var_t = table(*tabs[c])
            for li, (ss, ss2, cnt) in groups_exact(lens).items():
                if li >= n_bins or li - 1 >= len(mean_t):
                    continue
                m, v = Fraction(mean_t[li - 1]), Fraction(var_t[li - 1])
                mu += m * ss
                var += v * ss2
                mu_raw += m / li * cnt
                shared = True
        row = {"name": name, "chroms": chroms, "query": q,
               "tables": {c: {"kind": k, "len": K, "mean": table(k, K)[0], "var": table(k, K)[1]}
                          for c, (k, K) in tabs.items()},
               "observed": O, "record": shared}
        if shared:
            mu_f, var_f, mr_f = float(mu), float(var), float(mu_raw)
            llr = None
            if var > mu:
                row["branch"] = "nb"
                r = mu_f * mu_f / (var_f - mu_f)
                p = mu_f / var_f
                p_o = r / (r + O)
                if 0 < p_o < 1:
                    llr = float(nbinom.logpmf(O, r, p_o) - nbinom.logpmf(O, r, p))
            else:
                row["branch"] = "poisson"
                oc = Fraction(O) / (mu / mu_raw)
                assert oc.denominator == 1, name  # integer point for the pmf
                if oc > 0:
                    llr = float(poisson.logpmf(int(oc), int(oc)) - poisson.logpmf(int(oc), mr_f))
            row.update({"mu": mu_f, "var": var_f, "mu_raw": mr_f, "llr": llr})
            if llr is not None and llr > 0:
                tail = 0.5 * math.erfc(math.sqrt(llr))
                row["p"] = tail if O >= mu_f else 1.0 - tail
            else:
                row["p"] = 1.0
        rows.append(row)
    return rows


if __name__ == "__main__":
    for name, data in (("sweep_cases.json", sweep_cases()), ("moment_cases.json", moment_cases()),
                       ("stats_cases.json", stats_cases()), ("group_cases.json", group_cases()),
                       ("analytic_cases.json", analytic_cases())):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print("wrote", name)
