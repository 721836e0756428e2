"""Oracle and HIP path against the golden fixtures of tests/golden/, which
tests/golden/make_fixtures.py derives from first principles (pair brute
force, direct window sums in exact rationals, scipy log-pmf differences)
without the oracle or the reference's code."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def arrays(rows):
    a = np.array(rows, np.uint32).reshape(-1, 4)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), a[:, 3].copy()


SWEEP = load("sweep_cases.json")
MOMENTS = load("moment_cases.json")
STATS = load("stats_cases.json")


@pytest.mark.parametrize("k", range(len(SWEEP)))
def test_oracle_sweep_matches_fixture(orc, k):
    c = SWEEP[k]
    db, q = arrays(c["db"]), arrays(c["q"])
    for oc, obp in (orc.sweep_bruteforce(db, q, c["F"], c["nq"]),
                    orc.sweep_ranks(db, q, c["F"], c["nq"], c["nshards"])):
        assert oc.tolist() == c["counts"] and obp.tolist() == c["bp"]


@pytest.mark.parametrize("k", range(len(MOMENTS)))
def test_oracle_moments_match_fixture(orc, k):
    c = MOMENTS[k]
    iv = np.array(c["intervals"], np.uint32).reshape(-1, 2)
    G = orc.depth_exact(iv[:, 0], iv[:, 1], c["n_bins"])
    assert G.tolist() == c["G"]
    m, v, _, _ = orc.moments_exact(G, c["L"])
    np.testing.assert_allclose(m, c["mean"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(v, c["var"], rtol=1e-9, atol=1e-9)
    r = orc.moments_exact_fft(G, c["L"])
    if r is not None:
        np.testing.assert_allclose(r[0], c["mean"], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(r[1], c["var"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("k", range(len(STATS)))
def test_oracle_llr_and_tail_match_fixture(orc, k):
    c = STATS[k]
    if c["kind"] == "nb":
        llr = orc.nb_llr(c["mu"], c["var"], float(c["obs"]))
    else:
        llr = orc.poisson_llr(c["mu"], float(c["obs"]))
    assert llr is not None and llr == pytest.approx(c["llr"], rel=1e-9, abs=1e-9)
    # A&S 7.1.26 approximates erfc within 1.5e-7 absolute
    assert abs(0.5 * orc.erfc_as(math.sqrt(llr)) - c["tail"]) < 0.75e-7 + 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(len(SWEEP)))
def test_gpu_sweep_matches_fixture(chk, k):
    c = SWEEP[k]
    db, q = arrays(c["db"]), arrays(c["q"])
    ix = chk.Index(*db, c["nshards"], c["F"])
    for method in (1, 2):
        ix.tune("sweep.method", method)
        cnt, bp = ix.sweep(ix.query(q[0], q[1], q[2], q[3], c["nq"]))
        assert cnt.tolist() == c["counts"] and bp.tolist() == c["bp"], f"method {method}"


@pytest.mark.gpu
def test_gpu_moments_match_fixture(chk, orc):
    # every fixture track becomes one shard holding one DB file
    S = len(MOMENTS)
    sh, st, en = [], [], []
    for s, c in enumerate(MOMENTS):
        for a, e in c["intervals"]:
            sh.append(s)
            st.append(a)
            en.append(e)
    sizes = np.array([100 * c["n_bins"] for c in MOMENTS], np.uint64)
    ix = chk.Index(np.array(sh, np.uint32), np.array(st, np.uint32), np.array(en, np.uint32),
                   np.zeros(len(st), np.uint32), S, 1)
    ix.set_genome(sizes)
    ix.build_moments()
    so, sl, mean, var = ix.download_moments()
    checked = 0
    for s, c in enumerate(MOMENTS):
        if not c["intervals"]:
            assert so[s] < 0  # a file without intervals on the shard stores no table
            continue
        ls = orc.compact_ls(c["n_bins"])
        pos = {int(L): i for i, L in enumerate(ls)}
        for L, m, v in zip(c["L"], c["mean"], c["var"]):
            if L not in pos:
                continue
            i = so[s] + pos[L]
            assert mean[i] == pytest.approx(m, rel=1e-6, abs=1e-12), (c["name"], L)
            assert var[i] == pytest.approx(v, rel=1e-6, abs=1e-12), (c["name"], L)
            checked += 1
    assert checked > 50


GROUPS = load("group_cases.json")
ANALYTIC = load("analytic_cases.json")


@pytest.mark.parametrize("k", range(len(GROUPS)))
def test_oracle_group_lengths_match_fixture(orc, k):
    c = GROUPS[k]
    li, ss, ss2, cnt = orc.group_lengths(c["len_bp"])
    assert li.tolist() == c["l_int"] and cnt.tolist() == c["count"]
    np.testing.assert_allclose(ss, c["sum_scale"], rtol=1e-12)
    np.testing.assert_allclose(ss2, c["sum_scale2"], rtol=1e-12)
    # every interval lands in one group, and l_int * sum_scale restores sum l
    assert int(cnt.sum()) == len(c["len_bp"])
    assert float((li * ss).sum()) == pytest.approx(sum(c["len_bp"]) / 100.0, rel=1e-12)


def analytic_oracle_inputs(orc, c):
    """Groups (with each chromosome's table slot) and tables of one fixture cell."""
    tables, slot_of = [], {}
    for ch, t in c["tables"].items():
        slot_of[ch] = len(tables)
        tables.append(((c["chroms"][ch] + 99) // 100, np.array(t["mean"]), np.array(t["var"])))
    parts = [[], [], [], [], []]
    for ch, lens in c["query"].items():
        g = orc.group_lengths(lens)
        for a, v in zip(parts, g + (np.full(len(g[0]), slot_of.get(ch, -1), np.int32),)):
            a.append(v)
    return tuple(np.concatenate(a) for a in parts), tables


@pytest.mark.parametrize("k", range(len(ANALYTIC)))
def test_oracle_analytic_stats_match_fixture(orc, k):
    c = ANALYTIC[k]
    groups, tables = analytic_oracle_inputs(orc, c)
    got = orc.analytic_stats(groups, None, tables, float(c["observed"]))
    assert (got is not None) == c["record"], c["name"]
    if got is None:
        return
    p, llr = got
    if c["llr"] is None:
        assert llr is None and p == 1.0
    else:
        assert llr == pytest.approx(c["llr"], rel=1e-9, abs=1e-12), c["name"]
        # A&S 7.1.26 is within 1.5e-7 of erfc: p within 0.75e-7 of the fixture's
        assert abs(p - c["p"]) < 0.75e-7 + 1e-12, c["name"]


def analytic_gpu_layout(c):
    """DB rows of one fixture cell: intervals [0, take) per query chromosome
    that cover exactly `observed` bins of the query, whose intervals lie back
    to back from 0.  None when the cell cannot be laid out (O = 0 or more
    than the query's length)."""
    remaining = 100 * c["observed"]
    db = []
    for ch, lens in c["query"].items():
        take = min(sum(lens), remaining)
        if take > 0:
            db.append((ch, 0, take))
        remaining -= take
    if remaining > 0 or c["observed"] == 0:
        return None
    # every DB file also holds the chromosomes of its tables (the file "has"
    # those tables), far from the query
    for ch in c["tables"]:
        db.append((ch, 900_000, 900_100))
    return db


@pytest.mark.gpu
def test_gpu_analytic_stats_match_fixture(chk):
    cases = [c for c in ANALYTIC if analytic_gpu_layout(c) is not None]
    assert len(cases) >= 7
    names = list(cases[0]["chroms"])
    S, F = len(names), len(cases)
    sid = {ch: i for i, ch in enumerate(names)}
    dsh, dst, den, dfid = [], [], [], []
    qsh, qst, qen, qfile = [], [], [], []
    for f, c in enumerate(cases):
        for ch, a, e in analytic_gpu_layout(c):
            dsh.append(sid[ch]); dst.append(a); den.append(e); dfid.append(f)
        for ch, lens in c["query"].items():
            pos = 0
            for ln in lens:
                qsh.append(sid[ch]); qst.append(pos); qen.append(pos + ln); qfile.append(f)
                pos += ln
    u = lambda v: np.array(v, np.uint32)
    ix = chk.Index(u(dsh), u(dst), u(den), u(dfid), S, F)
    ix.set_genome(np.array([cases[0]["chroms"][n] for n in names], np.uint64))
    so = np.full(F * S, -1, np.int64)
    sl = np.zeros(F * S, np.uint32)
    mean, var = [], []
    off = 0
    for f, c in enumerate(cases):
        for ch, t in c["tables"].items():
            so[f * S + sid[ch]] = off
            sl[f * S + sid[ch]] = t["len"]
            mean += t["mean"]
            var += t["var"]
            off += t["len"]
    ix.upload_moments(so, sl, np.array(mean), np.array(var))
    out = ix.stats(ix.query(u(qsh), u(qst), u(qen), u(qfile), F))
    for f, c in enumerate(cases):
        assert int(out["bp"][f, f]) == 100 * c["observed"], c["name"]
        assert bool(out["flags"][f, f] & 1) == c["record"], c["name"]
        if not c["record"]:
            assert out["p"][f, f] == 1.0 and out["llr"][f, f] == 0.0
            continue
        assert bool(out["flags"][f, f] & 2) == (c["llr"] is not None), c["name"]
        if c["llr"] is not None:
            assert out["llr"][f, f] == pytest.approx(c["llr"], rel=1e-9, abs=1e-12), c["name"]
        assert abs(out["p"][f, f] - c["p"]) < 0.75e-7 + 1e-12, c["name"]
