"""Serving form: continuous batching over a resident index.  Every request's
row must equal the oracle's counts / bp, and be bit-identical (p-values too)
whether it ran alone or inside a mix of concurrent requests."""
import threading

import numpy as np
import pytest

from tests.cases import flatten, random_files

pytestmark = pytest.mark.gpu

S, SPAN = 3, 4_000_000
GENOME = np.array([4_100_000, 4_100_000, 4_100_000], np.uint64)


@pytest.fixture(scope="module")
def served(chk):
    rng = np.random.default_rng(41)
    F = 30
    db = flatten(random_files(rng, F, S, (300, 6000), SPAN, 20_000))
    ix = chk.Index(*db, S, F)
    ix.set_genome(GENOME)
    ix.build_moments()
    files = [random_files(rng, 1, S, n, SPAN, 20_000)[0]
             for n in (5, 60, 700, 4000, 25_000, 120_000, 2, 9000, 300, 50_000, 1, 15_000)]
    return ix, db, F, files


def oracle_row(orc, db, F, f):
    q = flatten([f])
    c, bp = orc.sweep_ranks(db, q, F, 1, S)
    return c[0], bp[0]


def test_alone_equals_in_mix(chk, orc, served):
    ix, db, F, files = served
    srv = chk.Server(ix, slots=2, max_batch_files=4, pvalues=True)
    a = files[3]
    alone = srv.wait(srv.submit(*a))
    oc, obp = oracle_row(orc, db, F, a)
    assert (alone["counts"].astype(np.int64) == oc).all()
    assert (alone["bp"].astype(np.int64) == obp).all()
    order = list(range(len(files)))
    order.insert(5, 3)
    order.insert(9, 3)
    tickets = [(i, srv.submit(*files[i])) for i in order]
    for i, t in tickets:
        r = srv.wait(t)
        oc, obp = oracle_row(orc, db, F, files[i])
        assert (r["counts"].astype(np.int64) == oc).all(), f"file {i}"
        assert (r["bp"].astype(np.int64) == obp).all(), f"file {i}"
        if i == 3:
            for k in ("counts", "bp", "p", "llr", "flags"):
                assert (r[k] == alone[k]).all(), k
    st = srv.stats()
    assert st["completed"] == len(tickets) + 1 and st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_tiny_budget_runs_requests_alone(chk, orc, served):
    ix, db, F, files = served
    srv = chk.Server(ix, slots=3, max_batch_files=64, hbm_budget=1)
    tickets = [(i, srv.submit(*files[i])) for i in range(6)]
    for i, t in tickets:
        r = srv.wait(t)
        oc, obp = oracle_row(orc, db, F, files[i])
        assert (r["counts"].astype(np.int64) == oc).all() and (r["bp"].astype(np.int64) == obp).all()
    assert srv.stats()["files_per_batch"] == 1.0
    srv.close()


def test_concurrent_submitters(chk, orc, served):
    ix, db, F, files = served
    srv = chk.Server(ix, slots=3, max_batch_files=8)
    results, errors = {}, []

    def client(k):
        try:
            for j in range(4):
                i = (k * 5 + j * 3) % len(files)
                results[(k, j)] = (i, srv.wait(srv.submit(*files[i])))
        except Exception as e:  # surfaced below
            errors.append(e)

    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for (k, j), (i, r) in results.items():
        oc, obp = oracle_row(orc, db, F, files[i])
        assert (r["counts"].astype(np.int64) == oc).all() and (r["bp"].astype(np.int64) == obp).all()
    srv.close()


def test_reserved_server_matches_alone(chk, orc, served):
    # slots grown up front for the largest batch (with p-values and their
    # length groups) give the same rows as a server that grows on demand
    ix, db, F, files = served
    plain = chk.Server(ix, slots=1, max_batch_files=4, pvalues=True)
    ref = [plain.wait(plain.submit(*f)) for f in files]
    plain.close()
    srv = chk.Server(ix, slots=3, max_batch_files=4, pvalues=True)
    srv.reserve(sum(len(f[1]) for f in files))
    tickets = [srv.submit(*f) for f in files]
    for i, t in enumerate(tickets):
        r = srv.wait(t)
        for k in ("counts", "bp", "p", "llr", "flags"):
            assert (r[k] == ref[i][k]).all(), (i, k)
    oc, obp = oracle_row(orc, db, F, files[5])
    assert (ref[5]["counts"].astype(np.int64) == oc).all() and (ref[5]["bp"].astype(np.int64) == obp).all()
    srv.close()


def test_large_requests_staged_over_host_threads(chk, orc, served):
    # batches above 2^18 intervals are copied into the slot's staging by
    # several host threads, each a range of positions that may cross request
    # boundaries (an empty request -- every interval on a shard the DB lacks --
    # included)
    ix, db, F, _ = served
    rng = np.random.default_rng(47)
    sizes = (300_000, 1, 280_000, 70_000)
    files = [random_files(rng, 1, S, n, SPAN, 20_000)[0] for n in sizes]
    absent = (np.full(500, S + 2, np.uint32), np.arange(500, dtype=np.uint32),
              np.arange(500, dtype=np.uint32) + 10)
    srv = chk.Server(ix, slots=1, max_batch_files=8)
    order = [files[0], absent, files[1], files[2], files[3]]
    tickets = [srv.submit(*f) for f in order]
    rows = [srv.wait(t) for t in tickets]
    assert (rows[1]["counts"] == 0).all() and (rows[1]["bp"] == 0).all()
    for r, f in zip(rows[:1] + rows[2:], files):
        oc, obp = oracle_row(orc, db, F, f)
        assert (r["counts"].astype(np.int64) == oc).all() and (r["bp"].astype(np.int64) == obp).all()
    srv.close()
