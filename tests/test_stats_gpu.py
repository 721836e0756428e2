"""GPU analytic p-values vs the oracle's FP64 restatement of
compute_analytic_stats (src/fourier/stats.rs:34-126) on the same exact moment
tables; plus the reference's stats unit cases (stats.rs:145-242) and the
self-enrichment integration test (tests/integration.rs:120-182).
Tolerance: 1e-3 absolute in -log10 p (observed 1e-9-level differences come
from FP64 summation order only)."""
import numpy as np
import pytest

from tests.cases import flatten, integration_setup, random_files
from tests.stats_cases import mlog10, oracle_cell_stats, oracle_tables

pytestmark = pytest.mark.gpu


def run_case(chk, orc, db, q, F, nq, S, sizes):
    ix = chk.Index(*db, S, F)
    ix.set_genome(sizes)
    so, sl, m, v, tabs = oracle_tables(orc, db, F, S, sizes)
    ix.upload_moments(so, sl, m, v)
    qq = ix.query(q[0], q[1], q[2], q[3] if nq > 1 else None, nq)
    out = ix.stats(qq)
    oc, obp = orc.sweep_ranks(db, q, F, nq, S)
    return out, oc, obp, tabs


def check_all(orc, out, oc, obp, q, F, nq, S, sizes, tabs, tol=1e-3):
    has_entry = {f for (f, _c) in tabs}
    n_emitted = 0
    for r in range(nq):
        for f in range(F):
            assert int(out["counts"][r, f]) == oc[r, f]
            exp = None
            if oc[r, f] > 0 and f in has_entry:
                exp = oracle_cell_stats(orc, q, r, f, S, sizes, tabs, obp[r, f] / 100.0)
            got = bool(out["flags"][r, f] & 1)
            assert got == (exp is not None), (r, f, exp)
            if exp is None:
                continue
            n_emitted += 1
            p_exp, llr_exp = exp
            assert bool(out["flags"][r, f] & 2) == (llr_exp is not None)
            assert abs(mlog10(out["p"][r, f]) - mlog10(p_exp)) < tol or \
                (out["p"][r, f] == 0 and p_exp == 0)
            if llr_exp is not None:
                assert out["llr"][r, f] == pytest.approx(llr_exp, rel=1e-9, abs=1e-9)
    return n_emitted


@pytest.mark.parametrize("seed", [0, 1])
def test_random_cells_match_oracle(chk, orc, seed):
    rng = np.random.default_rng(seed)
    F, nq, S = 6, 3, 4
    sizes = np.array([2_000_000, 1_500_000, 0, 800_000], np.uint64)  # shard 2 outside the genome
    db = flatten(random_files(rng, F, S, (50, 400), 700_000, 20_000))
    q = flatten(random_files(rng, nq, S, (20, 300), 700_000, 30_000))
    out, oc, obp, tabs = run_case(chk, orc, db, q, F, nq, S, sizes)
    assert check_all(orc, out, oc, obp, q, F, nq, S, sizes, tabs) > 5


def test_self_query_is_most_enriched(chk, orc):
    # tests/integration.rs:120-182: self p < other p and self p < 1e-10
    names24 = orc.CHROMS24
    sizes = np.array([30_000_000 if n in ("chr1", "chr2", "chr3") else 0 for n in names24], np.uint64)
    a = orc.generate_bed(77, 2000, 30_000_000)
    b = orc.generate_bed(88, 2000, 30_000_000)
    db = flatten([a, b])
    q = flatten([a])
    out, oc, obp, tabs = run_case(chk, orc, db, q, 2, 1, 24, sizes)
    assert out["flags"][0, 0] & 1 and out["flags"][0, 1] & 1
    assert out["p"][0, 0] < out["p"][0, 1]
    assert out["p"][0, 0] < 1e-10
    check_all(orc, out, oc, obp, q, 2, 1, 24, sizes, tabs)


def test_no_shared_chrom_no_record(chk, orc):
    # stats.rs:148-159: query on chr21, DB on chr22 -> None.  Here a DB file
    # whose only stored table is on another shard overlaps nothing either, so
    # use a shard outside the genome for the overlap instead.
    sizes = np.array([0, 1_000_000], np.uint64)
    db = (np.array([0, 1], np.uint32), np.array([100, 5000], np.uint32),
          np.array([900, 9000], np.uint32), np.array([0, 0], np.uint32))
    q = (np.array([0], np.uint32), np.array([0], np.uint32), np.array([1000], np.uint32),
         np.zeros(1, np.uint32))
    out, oc, obp, tabs = run_case(chk, orc, db, q, 1, 1, 2, sizes)
    assert out["counts"][0, 0] == 1
    assert not (out["flags"][0, 0] & 1)
    assert out["p"][0, 0] == 1.0 and out["llr"][0, 0] == 0.0  # absent record: p = 1


def test_fixed_moments_depletion(chk):
    # stats.rs:197-216: mean 20, var 100 moments; one 10-bin query interval;
    # observed 10 bins < mu = 20 -> 0.5 < p < 1
    S, F = 1, 1
    ix = chk.Index(np.zeros(1, np.uint32), np.array([0], np.uint32), np.array([1000], np.uint32),
                   np.zeros(1, np.uint32), S, F)
    ix.set_genome(np.array([50_818_468], np.uint64))
    n = 2000
    ix.upload_moments(np.array([0]), np.array([n], np.uint32), np.full(n, 20.0), np.full(n, 100.0))
    out = ix.stats(ix.query(np.zeros(1, np.uint32), np.array([0], np.uint32),
                            np.array([1000], np.uint32)))
    assert out["counts"][0, 0] == 1 and out["bp"][0, 0] == 1000
    assert out["flags"][0, 0] == 3
    assert 0.5 < out["p"][0, 0] < 1.0


def test_golden_dataset_stats(chk, orc):
    # genome table: chr1-chr3 (the other shards carry no moments)
    sizes_all = np.array([30_000_000] * 3 + [0] * 21, np.uint64)
    db, q = integration_setup(orc, 5, 1000, 2, 500, 30_000_000)
    out, oc, obp, tabs = run_case(chk, orc, db, q, 5, 2, 24, sizes_all)
    assert int((out["counts"] > 0).sum()) == 10 and int(out["counts"].sum()) == 58
    assert check_all(orc, out, oc, obp, q, 5, 2, 24, sizes_all, tabs) >= 5
