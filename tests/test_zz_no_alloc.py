"""chk_sweep_noalloc (one query file, static device scratch, no allocation)
against the allocating entry points chk_query_create + chk_sweep_async on the
same input, bit for bit, on two calls in a row (the second smaller, so stale
scratch contents would show), and against the oracle's rank sweep."""
import numpy as np
import pytest

from tests.cases import flatten, random_files

pytestmark = pytest.mark.gpu


def test_noalloc_matches_allocating_sweep(chk, orc):
    rng = np.random.default_rng(77)
    S, F = 3, 1500  # two DB file groups
    db = flatten(random_files(rng, F, S, (20, 200), 3_000_000, 40_000))
    ix = chk.Index(*db, S, F)
    dc, db_ = ix.alloc(8 * F), ix.alloc(8 * F)
    ac, ab = ix.alloc(8 * F), ix.alloc(8 * F)
    for n in (60_000, 3_000):
        q = random_files(rng, 1, S + 1, n, 3_000_000, 60_000)[0]  # shard id S is absent
        ix.sweep_noalloc(q[0], q[1], q[2], dc, db_)
        got_c, got_b = dc.to_numpy(np.uint64, F), db_.to_numpy(np.uint64, F)
        ix.sweep_async(ix.query(q[0], q[1], q[2]), ac, ab)
        ix.sync()
        exp_c, exp_b = ac.to_numpy(np.uint64, F), ab.to_numpy(np.uint64, F)
        assert (got_c == exp_c).all() and (got_b == exp_b).all(), n
        keep = q[0] < S
        oc, obp = orc.sweep_ranks(db, flatten([(q[0][keep], q[1][keep], q[2][keep])]), F, 1, S)
        assert (got_c.astype(np.int64) == oc[0]).all() and (got_b.astype(np.int64) == obp[0]).all()
        assert got_c.sum() > 0


def test_noalloc_rejects_oversized_and_invalid_input(chk):
    S, F = 1, 4
    z = np.zeros(8, np.uint32)
    ix = chk.Index(z[:4], np.arange(4, dtype=np.uint32) * 10, np.arange(4, dtype=np.uint32) * 10 + 5,
                   np.arange(4, dtype=np.uint32), S, F)
    dc, db_ = ix.alloc(8 * F), ix.alloc(8 * F)
    n = chk.MAX_NOALLOC + 1
    big = np.zeros(n, np.uint32)
    with pytest.raises(chk.ChuckleError, match="at most"):
        ix.sweep_noalloc(big, big, big + 1, dc, db_)
    with pytest.raises(chk.ChuckleError, match="less than end"):
        ix.sweep_noalloc(z[:1], np.array([9], np.uint32), np.array([9], np.uint32), dc, db_)
    ix.sweep_noalloc(z[:0], z[:0], z[:0], dc, db_)  # an empty file gives zeros
    assert (dc.to_numpy(np.uint64, F) == 0).all() and (db_.to_numpy(np.uint64, F) == 0).all()
