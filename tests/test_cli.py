"""`chuckle add` / `info`: the reference's on-disk format (meta.json,
layer_K.{pos,sid,idx}; src/io/meta.rs, src/io/layer.rs, src/io/jump.rs,
src/tasks/build.rs:86-276) and BED parsing rules (src/io/parse.rs:53-141),
checked against the oracle's layer assignment and jump tables.  `add` sorts
its buckets on the device, so those tests are marked gpu; parse errors are
reported before the device is touched and run on the CPU."""
import gzip
import json
import os

import numpy as np
import pytest

from tests.cli_cases import CLI, run, scenario

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="chuckle CLI not built")


def read_layer(db, shard, k):
    base = os.path.join(db, "shards", shard, f"layer_{k}")
    pos = np.fromfile(base + ".pos", np.uint32).reshape(-1, 2)
    sid = np.fromfile(base + ".sid", np.uint32)
    idx = np.fromfile(base + ".idx", np.uint64)
    return pos, sid, idx


def check_db(orc, db, expected):
    """expected: list of (shard name, start, end, sid) of every interval."""
    meta = json.load(open(os.path.join(db, "meta.json")))
    assert meta["layer_config"] == {"min_size": 128, "growth_factor": 4}
    got = []
    for shard in meta["shards"]:
        for k in range(meta["num_layers"]):
            if not os.path.exists(os.path.join(db, "shards", shard, f"layer_{k}.pos")):
                continue
            pos, sid, idx = read_layer(db, shard, k)
            assert len(pos) == len(sid)
            assert (np.diff(pos[:, 0].astype(np.int64)) >= 0).all(), "layer not sorted by start"
            for s, e in pos:
                assert orc.layer_for_size(int(e - s)) == k
            tile = max(1, orc.layer_max_size(k) // 2)
            assert idx.tolist() == orc.build_jump_table(pos[:, 0], tile).tolist()
            got += [(shard, int(s), int(e), int(d)) for (s, e), d in zip(pos, sid)]
    assert sorted(got) == sorted(expected)
    return meta


@pytest.mark.gpu
def test_add_writes_reference_layers(orc, tmp_path):
    src, _ = scenario(orc, str(tmp_path), 4, 700, 0, 0, 5_000_000)
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db)
    meta = json.load(open(os.path.join(db, "meta.json")))
    names = {int(k): v["name"] for k, v in meta["sid_map"].items()}
    assert sorted(names.values()) == [f"source_{i}.bed" for i in range(4)]
    expected = []
    for sid, name in names.items():
        i = int(name.split("_")[1].split(".")[0])
        c, s, e = orc.generate_bed(100 + i, 700, 5_000_000)
        expected += [(orc.CHROMS24[a], int(b), int(d), sid) for a, b, d in zip(c, s, e)]
    check_db(orc, db, expected)


@pytest.mark.gpu
def test_add_appends_and_merges(orc, tmp_path):
    src, _ = scenario(orc, str(tmp_path), 3, 500, 0, 0, 2_000_000)
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db, "--batch-size", "1")
    extra = tmp_path / "extra"
    extra.mkdir()
    orc.write_bed(str(extra / "late.bed"), *orc.generate_bed(999, 400, 2_000_000))
    run("add", "-i", str(extra), "-d", db)
    meta = json.load(open(os.path.join(db, "meta.json")))
    names = {int(k): v["name"] for k, v in meta["sid_map"].items()}
    assert names[3] == "late.bed" and len(names) == 4
    expected = []
    for sid, name in names.items():
        seed = 999 if name == "late.bed" else 100 + int(name.split("_")[1].split(".")[0])
        n = 400 if name == "late.bed" else 500
        c, s, e = orc.generate_bed(seed, n, 2_000_000)
        expected += [(orc.CHROMS24[a], int(b), int(d), sid) for a, b, d in zip(c, s, e)]
    check_db(orc, db, expected)


@pytest.mark.gpu
def test_bed_parsing_rules(orc, tmp_path):
    src = tmp_path / "src"
    src.mkdir()
    (src / "a.bed").write_text("# comment\n\nchr1\t10\t20\n30 40\n50\t60\tchr2\tname\t7\n"
                               "chrX 5 9 extra 1 2\n")
    with gzip.open(src / "b.bed.gz", "wt") as f:
        f.write("chr1\t100\t200\n")
    (src / "ignored.txt").write_text("chr1 1 2\n")
    db = str(tmp_path / "db")
    run("add", "-i", str(src), "-d", db)
    meta = json.load(open(os.path.join(db, "meta.json")))
    names = {v["name"]: int(k) for k, v in meta["sid_map"].items()}
    assert set(names) == {"a.bed", "b.bed.gz"}
    a, b = names["a.bed"], names["b.bed.gz"]
    check_db(orc, db, [("chr1", 10, 20, a), ("default", 30, 40, a), ("chr2", 50, 60, a),
                       ("chrX", 5, 9, a), ("chr1", 100, 200, b)])


def test_parse_error_line_number_in_a_large_file(tmp_path):
    # a file of about 30 MB is parsed in several chunks over host threads:
    # the error's line number must count every line of the earlier chunks
    src = tmp_path / "src"
    src.mkdir()
    n_before = 1_200_000
    rng = np.random.default_rng(3)
    s = rng.integers(0, 10**8, n_before)
    lines = [f"chr{1 + i % 22}\t{a}\t{a + 100}" for i, a in enumerate(s.tolist())]
    lines[1000] = "# a comment line"
    lines[2000] = ""
    body = "\n".join(lines) + "\nchr1\t500\t400\n" + "chr1\t1\t2\n" * 10
    (src / "big.bed").write_text(body)
    assert len(body) > 2**24
    r = run("add", "-i", str(src), "-d", str(tmp_path / "db"), check=False)
    assert r.returncode == 1
    assert r.stderr.strip() == f"Error: Line {n_before + 1}: start (500) must be less than end (400)"


def test_bed_errors_exit_1(tmp_path):
    src = tmp_path / "src"
    src.mkdir()
    (src / "bad.bed").write_text("chr1 10 20\nchr1 30 30\n")
    r = run("add", "-i", str(src), "-d", str(tmp_path / "db"), check=False)
    assert r.returncode == 1
    assert r.stderr.strip() == "Error: Line 2: start (30) must be less than end (30)"
    (src / "bad.bed").write_text("chr1 10\n")
    r = run("add", "-i", str(src), "-d", str(tmp_path / "db2"), check=False)
    assert r.returncode == 1 and "Line 1: invalid format" in r.stderr


@pytest.mark.gpu
def test_info(orc, tmp_path):
    src, _ = scenario(orc, str(tmp_path), 2, 100, 0, 0, 1_000_000)
    db = str(tmp_path / "db")
    run("add", "-i", src, "-d", db)
    out = run("info", "-d", db).stdout.splitlines()
    assert out[0] == f"Database: {db}"
    assert out[1] == "Layer config: min_size=128, growth_factor=4"
    assert out[2].startswith("Layers used: ")
    assert out[3].startswith("Shards (")
    assert out[5] == "Sources (2):"
    assert sorted(out[6:]) == ["  SID    0: source_0.bed", "  SID    1: source_1.bed"] or \
        sorted(x.split(": ")[1] for x in out[6:]) == ["source_0.bed", "source_1.bed"]


def _read_trials(d):
    out = []
    for name in sorted(os.listdir(d)):
        rows = []
        with open(os.path.join(d, name)) as f:
            for line in f:
                c, s, e = line.split("\t")
                rows.append((c, int(s), int(e)))
        out.append(rows)
    return out


@pytest.mark.parametrize("mode,wl", [("circular", False), ("shuffle", False), ("circular", True)])
def test_montecarlo_trials_match_oracle(orc, tmp_path, mode, wl):
    # src/bin/montecarlo.rs:79-163, 232-311: same seed -> the same shifted /
    # shuffled trial files, batch by batch (chromosomes in name order, one not
    # in hg38, intervals reaching past both chromosome ends, a whitelist)
    q = tmp_path / "q.bed"
    rows = [("chr2", 5, 400), ("chr1", 248_956_000, 248_956_422), ("chr1", 10, 20_000),
            ("chrUn", 100, 200), ("chr21", 46_709_000, 46_709_983), ("chr10", 1_000_000, 1_300_000),
            ("chr1", 100_000_000, 100_000_050), ("chrY", 57_000_000, 57_227_415)]
    q.write_text("".join(f"{c}\t{s}\t{e}\n" for c, s, e in rows))
    qmap = {}
    for c, s, e in rows:
        qmap.setdefault(c, []).append((s, e))
    args = ["montecarlo", "-d", str(tmp_path / "nodb"), "-q", str(q), "-t", "7,20,12", "-b", "5",
            "-s", "1234567", "--mode", mode, "--keep-trials", str(tmp_path / "keep"), "--dry-run"]
    wmap = None
    if wl:
        w = tmp_path / "wl.bed"
        wrows = [("chr1", 0, 1_000_000), ("chr1", 50_000_000, 150_000_000), ("chr2", 10_000, 90_000_000),
                 ("chr10", 0, 20_000_000), ("chr21", 1_000_000, 46_000_000)]
        w.write_text("".join(f"{c}\t{s}\t{e}\n" for c, s, e in wrows))
        args += ["-w", str(w)]
        wmap = {}
        for c, s, e in wrows:
            wmap.setdefault(c, []).append((s, e))
    run(*args)
    batches = list(orc.montecarlo_trials(qmap, 1234567, [7, 20, 12], 5, mode == "shuffle", wmap))
    assert [b[0] for b in batches] == [5, 2, 5, 5, 3]  # milestones 7, 12, 20 in batches of <= 5
    for k, (_, trials) in enumerate(batches):
        d = tmp_path / "keep" / f"batch_{k}"
        got = _read_trials(d) if d.exists() else []
        assert got == trials, k
    assert sum(len(t) for _, b in batches for t in b) > 0


HG38_SIZES = {"chr1": 248_956_422, "chr2": 242_193_529, "chr10": 133_797_422, "chr21": 46_709_983,
              "chrY": 57_227_415}


def _mc_dry(tmp_path, rows, mode, trials=60, wl_rows=None):
    q = tmp_path / "q.bed"
    q.write_text("".join(f"{c}\t{s}\t{e}\n" for c, s, e in rows))
    args = ["montecarlo", "-d", str(tmp_path / "nodb"), "-q", str(q), "-t", str(trials), "-b", "16",
            "-s", "99", "--mode", mode, "--keep-trials", str(tmp_path / "keep"), "--dry-run"]
    if wl_rows:
        w = tmp_path / "wl.bed"
        w.write_text("".join(f"{c}\t{s}\t{e}\n" for c, s, e in wl_rows))
        args += ["-w", str(w)]
    run(*args)
    out = []
    for b in sorted(os.listdir(tmp_path / "keep")):
        out += _read_trials(tmp_path / "keep" / b)
    return out


def test_montecarlo_circular_shift_properties(tmp_path):
    # properties of src/bin/montecarlo.rs:79-122 that hold whatever the draws:
    # a trial is the query moved by ONE shift on every chromosome: each
    # original interval whose image meets [0, size) appears once, clipped to
    # it (so unclipped ones keep their length and their gaps), the others
    # vanish; chromosomes outside hg38 never appear
    rows = [("chr1", 1_000, 5_000), ("chr1", 20_000, 20_100), ("chr1", 7_000_000, 7_500_000),
            ("chr21", 3_000, 4_000), ("chr21", 40_000_000, 40_000_010), ("chrUn", 5, 50),
            ("chrY", 1_000_000, 1_000_300)]
    trials = _mc_dry(tmp_path, rows, "circular")
    assert len(trials) > 30
    orig = [(c, a, b) for c, a, b in rows if c in HG38_SIZES]

    def image(sh):
        out = []
        for c, a, b in orig:
            lo, hi = max(0, a + sh), min(HG38_SIZES[c], b + sh)
            if lo < hi:
                out.append((c, lo, hi))
        return sorted(out)

    for t in trials:
        c0, s0, e0 = t[0]
        cands = {s0 - a for c, a, b in orig if c == c0} | {e0 - b for c, a, b in orig if c == c0}
        assert any(image(sh) == sorted(t) for sh in cands), t


def test_montecarlo_shuffle_properties(tmp_path):
    # src/bin/montecarlo.rs:124-163: every interval no longer than its
    # chromosome is placed once per trial, with its length, inside
    # [0, size); chromosomes outside hg38 are dropped
    rows = [("chr2", 10, 110), ("chr2", 500, 50_500), ("chr2", 9, 10), ("chr10", 0, 1_000_000),
            ("chrUn", 1, 2), ("chr21", 46_000_000, 46_709_983)]
    trials = _mc_dry(tmp_path, rows, "shuffle")
    want = sorted((c, e - s) for c, s, e in rows if c in HG38_SIZES)
    assert len(trials) == 60
    for t in trials:
        assert sorted((c, e - s) for c, s, e in t) == want
        for c, s, e in t:
            assert 0 <= s < e <= HG38_SIZES[c]
    # placements vary from trial to trial
    assert len({tuple(t) for t in trials}) == 60


def test_montecarlo_whitelist_clips_into_regions(tmp_path):
    # with a whitelist every trial interval lies inside one whitelist region
    # of its chromosome (src/bin/montecarlo.rs:71-77)
    rows = [("chr1", 1_000, 90_000), ("chr1", 5_000_000, 5_100_000), ("chr2", 0, 10_000)]
    wl = [("chr1", 0, 2_000_000), ("chr1", 60_000_000, 61_000_000), ("chr2", 100, 200_000_000)]
    for mode in ("circular", "shuffle"):
        (tmp_path / mode).mkdir()
        trials = _mc_dry(tmp_path / mode, rows, mode, wl_rows=wl)
        assert trials
        for t in trials:
            for c, s, e in t:
                assert any(c == wc and ws <= s < e <= we for wc, ws, we in wl), (mode, c, s, e)
