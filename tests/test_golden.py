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
