"""Batched form: one HIP-graph capture of a many-file sweep (+ p-values),
replayed; and the autotuner, which may change only the speed, never the
result.  Results are checked against the oracle (counts, bp) and against
the eager path (p-values)."""
import numpy as np
import pytest

from tests.cases import flatten, random_files

pytestmark = pytest.mark.gpu


def mixed_batch(rng, S, span, maxlen, sizes):
    return flatten([random_files(rng, 1, S, n, span, maxlen)[0] for n in sizes])


def test_graph_replays_match_eager_and_oracle(chk, orc):
    rng = np.random.default_rng(31)
    F, S, span = 40, 3, 5_000_000
    db = flatten(random_files(rng, F, S, (300, 5000), span, 20_000))
    sizes = [1, 17, 250, 3000, 12_000, 40_000, 900, 5]
    q = mixed_batch(rng, S, span, 20_000, sizes)
    B = len(sizes)
    ix = chk.Index(*db, S, F)
    ix.tune("sweep.sparse_max", 1000)  # both paths inside the graph
    qq = ix.query(q[0], q[1], q[2], q[3], B)
    oc, obp = orc.sweep_ranks(db, q, F, B, S)
    dc, dbp = ix.alloc(8 * B * F), ix.alloc(8 * B * F)
    g = ix.graph(qq, dc, dbp)
    for rep in range(3):
        # eager work of the program between replays
        ec, eb = ix.sweep(qq)
        assert (ec.astype(np.int64) == oc).all() and (eb.astype(np.int64) == obp).all()
        g.launch()
        ix.sync()
        c = dc.to_numpy(np.uint64, (B, F))
        bp = dbp.to_numpy(np.uint64, (B, F))
        assert (c.astype(np.int64) == oc).all(), f"replay {rep}"
        assert (bp.astype(np.int64) == obp).all(), f"replay {rep}"
    g.close()


def test_graph_with_pvalues(chk, orc):
    rng = np.random.default_rng(32)
    F, S = 12, 3
    sizes_bp = [3_000_000, 2_000_000, 1_500_000]
    db = flatten(random_files(rng, F, S, (500, 3000), 1_400_000, 5_000))
    sizes = [40, 2000, 300]
    q = mixed_batch(rng, S, 1_400_000, 5_000, sizes)
    B = len(sizes)
    ix = chk.Index(*db, S, F)
    ix.set_genome(np.array(sizes_bp, np.uint64))
    ix.build_moments()
    ix.tune("sweep.sparse_max", 500)
    qq = ix.query(q[0], q[1], q[2], q[3], B)
    ref = ix.stats(qq)
    cells = B * F
    dc, dbp = ix.alloc(8 * cells), ix.alloc(8 * cells)
    dp, dl, dfl = ix.alloc(8 * cells), ix.alloc(8 * cells), ix.alloc(cells)
    g = ix.graph(qq, dc, dbp, dp, dl, dfl)
    for rep in range(3):
        again = ix.stats(qq)
        assert (again["p"] == ref["p"]).all()
        g.launch()
        ix.sync()
        assert (dc.to_numpy(np.uint64, (B, F)) == ref["counts"]).all(), f"replay {rep}"
        assert (dp.to_numpy(np.float64, (B, F)) == ref["p"]).all(), f"replay {rep}"
        assert (dl.to_numpy(np.float64, (B, F)) == ref["llr"]).all(), f"replay {rep}"
        assert (dfl.to_numpy(np.uint8, (B, F)) == ref["flags"]).all(), f"replay {rep}"
    g.close()


def test_graph_refuses_after_retune(chk):
    rng = np.random.default_rng(33)
    db = flatten(random_files(rng, 3, 1, 200, 100_000, 500))
    ix = chk.Index(*db, 1, 3)
    qq = ix.query(db[0][:50], db[1][:50], db[2][:50])
    dc, dbp = ix.alloc(8 * 3), ix.alloc(8 * 3)
    g = ix.graph(qq, dc, dbp)
    g.launch()
    ix.tune("sweep.chunk", 3)
    with pytest.raises(chk.ChuckleError):
        g.launch()
    g.close()


def test_autotune_changes_speed_not_results(chk, orc):
    rng = np.random.default_rng(34)
    F, S, span = 25, 4, 8_000_000
    db = flatten(random_files(rng, F, S, (2000, 20_000), span, 30_000))
    sizes = [3, 80, 700, 6000, 30_000, 120_000]
    q = mixed_batch(rng, S, span, 30_000, sizes)
    B = len(sizes)
    ix = chk.Index(*db, S, F)
    qq = ix.query(q[0], q[1], q[2], q[3], B)
    dc, dbp = ix.alloc(8 * B * F), ix.alloc(8 * B * F)
    best, knobs, report = ix.autotune(qq, dc, dbp, reps=2)
    assert best > 0 and len(report) >= 5
    # with dense files in the kept plan the dense launch shape is searched
    # too (it once went unsearched when the last threshold candidate had sent
    # every file sparse)
    if qq.info()["dense_files"]:
        assert any(r.startswith("sweep.items=") for r in report), report
    c, bp = ix.sweep(qq)
    oc, obp = orc.sweep_ranks(db, q, F, B, S)
    assert (c.astype(np.int64) == oc).all() and (bp.astype(np.int64) == obp).all()
