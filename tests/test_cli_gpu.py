"""`chuckle add` + `chuckle query` end to end on the GPU, mirroring the
reference's integration tests (tests/integration.rs:57-182) on the same
seeded BED files: totals of hits and bp, nonzero pairs, and the self-query
enrichment; every (query, DB) record is also checked against the oracle's
brute force by name."""
import math
import os

import numpy as np
import pytest

from tests.cli_cases import CLI, results, run, scenario

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")]


def oracle_by_name(orc, n_db, db_iv, q_names, q_iv, gsize):
    """{(query file, db file): (count, bp)} by brute force."""
    out = {}
    for qn in q_names:
        qi = int(qn.split("_")[1].split(".")[0])
        qc, qs, qe = orc.generate_bed(200 + qi, q_iv, gsize)
        for d in range(n_db):
            dc, ds, de = orc.generate_bed(100 + d, db_iv, gsize)
            db = (dc.astype(np.uint32), ds, de, np.zeros(len(ds), np.uint32))
            q = (qc.astype(np.uint32), qs, qe, np.zeros(len(qs), np.uint32))
            c, bp = orc.sweep_bruteforce(db, q, 1, 1)
            out[(qn, f"source_{d}.bed")] = (int(c[0, 0]), int(bp[0, 0]))
    return out


def test_counts_nonzero_pairs(orc, tmp_path):
    src, qdir = scenario(orc, str(tmp_path), 5, 1000, 2, 500, 30_000_000)
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db)
    out = str(tmp_path / "out.json")
    r = run("query", "-d", db, "-q", qdir, "-o", out)
    assert "Query complete: 2 query files" in r.stdout
    res = results(out)
    assert len(res) == 10 and sum(x["overlap_count"] for x in res) == 58  # integration.rs:110-114
    assert all("p_value" not in x for x in res)
    counts = [x["overlap_count"] for x in res]
    assert counts == sorted(counts, reverse=True)
    ora = oracle_by_name(orc, 5, 1000, ["query_0.bed", "query_1.bed"], 500, 30_000_000)
    for x in res:
        assert x["overlap_count"] == ora[(x["query_name"], x["db_name"])][0]


def test_stats_totals_match_intervals_golden(orc, tmp_path):
    src, qdir = scenario(orc, str(tmp_path), 10, 2000, 1, 1000, 50_000_000)
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db, "--stats")
    assert os.path.exists(os.path.join(db, "momentmap.chk"))
    out = str(tmp_path / "out.json")
    run("query", "-d", db, "-q", qdir, "-o", out, "--stats")
    res = results(out)
    # integration.rs:60-75: 158 hits, 413,652 bp over all pairs
    assert sum(x["overlap_count"] for x in res) == 158
    assert round(sum(x["observed_bins"] for x in res) * 100) == 413_652
    ps = [x["p_value"] for x in res]
    assert ps == sorted(ps) and all(0.0 <= p <= 1.0 for p in ps)
    ora = oracle_by_name(orc, 10, 2000, ["query_0.bed"], 1000, 50_000_000)
    for x in res:
        c, bp = ora[(x["query_name"], x["db_name"])]
        assert x["overlap_count"] == c and x["observed_bins"] == bp / 100.0
    # the same DB queried without the stored tables (rebuilt on the device)
    os.remove(os.path.join(db, "momentmap.chk"))
    out2 = str(tmp_path / "out2.json")
    run("query", "-d", db, "-q", qdir, "-o", out2, "--stats")
    assert results(out2) == res


def test_self_query_is_most_enriched(orc, tmp_path):
    src = tmp_path / "src"
    src.mkdir()
    orc.write_bed(str(src / "self.bed"), *orc.generate_bed(77, 2000, 30_000_000))
    orc.write_bed(str(src / "other.bed"), *orc.generate_bed(88, 2000, 30_000_000))
    db = str(tmp_path / "db")
    run("add", "-i", str(src), "-d", db, "--stats")
    out = str(tmp_path / "out.json")
    run("query", "-d", db, "-q", str(src / "self.bed"), "-o", out, "--stats")
    res = {x["db_name"]: x for x in results(out)}
    assert res["self.bed"]["p_value"] < res["other.bed"]["p_value"]  # integration.rs:177-181
    assert res["self.bed"]["p_value"] < 1e-10
    assert results(out)[0]["db_name"] == "self.bed"


def test_query_file_with_unknown_shards_and_empty_file(orc, tmp_path):
    src, qdir = scenario(orc, str(tmp_path), 2, 300, 1, 200, 5_000_000)
    with open(os.path.join(qdir, "query_9.bed"), "w") as f:
        f.write("# nothing here\n")
    with open(os.path.join(qdir, "query_5.bed"), "w") as f:
        f.write("chrUn 1 1000000\nchr1 0 5000000\n")
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db)
    out = str(tmp_path / "out.json")
    r = run("query", "-d", db, "-q", qdir, "-o", out)
    assert "Query complete: 2 query files" in r.stdout  # the empty file takes no row
    names = {x["query_name"] for x in results(out)}
    assert "query_9.bed" not in names


def _rust_exp4(v):
    m, e = f"{v:.4e}".split("e")
    return f"{m}e{int(e)}"


@pytest.mark.parametrize("mode", ["circular", "shuffle"])
def test_montecarlo_pvalues_match_oracle(orc, tmp_path, mode):
    # src/bin/montecarlo.rs:187-356: observed counts, trial batches, exceed
    # counts (count >= observed; every trial for observed = 0) and the
    # milestone tables, all recomputed with the oracle's sweep
    src, qdir = scenario(orc, str(tmp_path), 4, 3000, 1, 2000, 30_000_000)
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db)
    q = os.path.join(qdir, "query_0.bed")
    r = run("montecarlo", "-d", db, "-q", q, "-t", "6,15", "-b", "4", "-s", "77", "--mode", mode)
    import json
    meta = json.load(open(os.path.join(db, "meta.json")))
    names = {int(k): v["name"] for k, v in meta["sid_map"].items()}
    chroms = {c: i for i, c in enumerate(orc.CHROMS24)}
    dbf = {}
    for d, nm in names.items():
        c, s, e = orc.generate_bed(100 + int(nm.split("_")[1].split(".")[0]), 3000, 30_000_000)
        dbf[d] = (c.astype(np.uint32), s, e)
    F = max(names) + 1
    db_arr = tuple(np.concatenate([dbf[d][k] for d in sorted(dbf)]) for k in range(3))
    db_arr = db_arr + (np.concatenate([np.full(len(dbf[d][1]), d, np.uint32) for d in sorted(dbf)]),)

    def counts(files):
        sh, st, en, qf = [], [], [], []
        for i, rows in enumerate(files):
            for c, s, e in rows:
                sh.append(chroms.get(c, 99))
                st.append(s)
                en.append(e)
                qf.append(i)
        qa = tuple(np.array(v, np.uint32) for v in (sh, st, en, qf))
        return orc.sweep_ranks(db_arr, qa, F, len(files), 100)[0]

    qc, qs, qe = orc.generate_bed(200, 2000, 30_000_000)
    qmap = {}
    for c, s, e in zip(qc.tolist(), qs.tolist(), qe.tolist()):
        qmap.setdefault(orc.CHROMS24[c], []).append((s, e))
    obs = counts([[(orc.CHROMS24[c], s, e) for c, s, e in zip(qc.tolist(), qs.tolist(), qe.tolist())]])[0]
    exceed = np.zeros(F, np.int64)
    total, lines = 0, []
    ms = [6, 15]
    for to_run, trials in orc.montecarlo_trials(qmap, 77, ms, 4, mode == "shuffle"):
        exceed[obs == 0] += to_run
        if trials:
            cnt = counts(trials)
            exceed += ((cnt >= obs[None, :]) & (obs[None, :] > 0)).sum(axis=0)
        total += to_run
        for m in ms:
            if total >= m and all(f"# milestone={m} " not in x for x in lines):
                lines.append(f"# milestone={m} trials={total}")
                lines.append("db_name\tobserved\tp_value")
                for d in sorted(names):
                    p = 1.0 if obs[d] == 0 else exceed[d] / total
                    lines.append(f"{names[d]}\t{obs[d]}\t{_rust_exp4(p)}")
                lines.append("")
    assert r.stdout.split("\n")[:len(lines)] == lines
    assert any(obs > 0)


def test_index_open_and_query_open_match_oracle(chk, orc, tmp_path):
    # a CLI-written database opened through the C ABI by path, queried with
    # files parsed by the library, against the oracle by name
    import ctypes as C
    src, qdir = scenario(orc, str(tmp_path), 6, 1500, 3, 800, 20_000_000)
    (tmp_path / "q" / "empty.bed").write_text("# nothing\n")
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db, "--sorted-sids")
    L = chk.lib()
    L.chk_index_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
    L.chk_index_file_name.argtypes = [C.c_void_p, C.c_uint32, C.c_char_p, C.c_uint64, C.c_void_p]
    L.chk_index_shard_name.argtypes = [C.c_void_p, C.c_uint32, C.c_char_p, C.c_uint64, C.c_void_p]
    L.chk_query_open.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_char_p), C.c_void_p,
                                 C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    assert L.chk_index_open(db.encode(), 0, C.byref(h)) == 0, L.chk_last_error()
    S, F = C.c_uint32(), C.c_uint32()
    assert L.chk_index_info(h, C.byref(S), C.byref(F), None, None) == 0
    buf = C.create_string_buffer(256)
    names = []
    for f in range(F.value):
        assert L.chk_index_file_name(h, f, buf, 256, None) == 0
        names.append(buf.value.decode())
    assert names == [f"source_{i}.bed" for i in range(6)]  # --sorted-sids: name order
    shards = []
    for s_ in range(S.value):
        assert L.chk_index_shard_name(h, s_, buf, 256, None) == 0
        shards.append(buf.value.decode())
    qfiles = sorted(os.path.join(qdir, x) for x in os.listdir(qdir))
    arr = (C.c_char_p * len(qfiles))(*[x.encode() for x in qfiles])
    rows = np.zeros(len(qfiles), np.uint32)
    q = C.c_void_p()
    assert L.chk_query_open(h, len(qfiles), arr, rows.ctypes.data_as(C.c_void_p), C.byref(q)) == 0, \
        L.chk_last_error()
    assert rows[[os.path.basename(x) == "empty.bed" for x in qfiles]].tolist() == [0xFFFFFFFF]
    nq = int((rows != 0xFFFFFFFF).sum())
    cnt = np.zeros((nq, F.value), np.uint64)
    bp = np.zeros((nq, F.value), np.uint64)
    assert L.chk_sweep(h, q, cnt.ctypes.data_as(C.c_void_p), bp.ctypes.data_as(C.c_void_p)) == 0
    L.chk_query_free(q)
    L.chk_index_free(h)
    for k, path in enumerate(qfiles):
        if rows[k] == 0xFFFFFFFF:
            continue
        qi = int(os.path.basename(path).split("_")[1].split(".")[0])
        qc, qs, qe = orc.generate_bed(200 + qi, 800, 20_000_000)
        for f in range(F.value):
            dc, ds, de = orc.generate_bed(100 + f, 1500, 20_000_000)
            db_ = (dc.astype(np.uint32), ds, de, np.zeros(len(ds), np.uint32))
            c, b = orc.sweep_bruteforce(db_, (qc.astype(np.uint32), qs, qe, np.zeros(len(qs), np.uint32)),
                                        1, 1)
            assert int(cnt[rows[k], f]) == c[0, 0] and int(bp[rows[k], f]) == b[0, 0]


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def _counts_by_pair(path):
    return {(x["query_name"], x["db_name"]): x["overlap_count"] for x in results(path)}


# src/tasks/query/mod.rs:455-547: the same DB and query files as the reference's
# batching tests; half-open overlaps by hand (touching intervals do not count)
BATCH_DB = {"a.bed": "chr1\t100\t200\nchr1\t300\t400\n", "b.bed": "chr1\t150\t250\n",
            "c.bed": "chr1\t350\t500\n"}
BATCH_EXPECTED = {("q1.bed", "a.bed"): 1, ("q1.bed", "b.bed"): 1, ("q2.bed", "a.bed"): 1,
                  ("q2.bed", "b.bed"): 1, ("q3.bed", "a.bed"): 1, ("q3.bed", "c.bed"): 1,
                  ("q4.bed", "a.bed"): 2, ("q4.bed", "b.bed"): 1, ("q4.bed", "c.bed"): 1}


def test_batch_query_matches_unbatched(tmp_path):
    src, qdir = tmp_path / "src", tmp_path / "q"
    src.mkdir()
    qdir.mkdir()
    for name, text in BATCH_DB.items():
        _write(src / name, text)
    for name, text in (("q1.bed", "chr1\t100\t200\n"), ("q2.bed", "chr1\t200\t350\n"),
                       ("q3.bed", "chr1\t300\t450\n"), ("q4.bed", "chr1\t50\t600\n")):
        _write(qdir / name, text)
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db)
    got = {}
    for bs in (0, 1, 3):
        out = str(tmp_path / f"out_{bs}.json")
        run("query", "-d", db, "-q", qdir, "-o", out, *(["--batch-size", bs] if bs else []))
        got[bs] = _counts_by_pair(out)
    assert got[0] == BATCH_EXPECTED
    assert got[1] == got[0] and got[3] == got[0]


def test_batch_add_matches_unbatched(tmp_path):
    src, qdir = tmp_path / "src", tmp_path / "q"
    src.mkdir()
    qdir.mkdir()
    for name, text in BATCH_DB.items():
        _write(src / name, text)
    _write(qdir / "q.bed", "chr1\t100\t500\n")
    got = {}
    for bs in (0, 1):
        db = str(tmp_path / f"db_{bs}")
        run("add", "-i", src, "-d", db, *(["--batch-size", bs] if bs else []))
        out = str(tmp_path / f"out_{bs}.json")
        run("query", "-d", db, "-q", qdir / "q.bed", "-o", out)
        got[bs] = _counts_by_pair(out)
    assert got[0] == {("q.bed", "a.bed"): 2, ("q.bed", "b.bed"): 1, ("q.bed", "c.bed"): 1}
    assert got[1] == got[0]
