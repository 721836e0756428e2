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
