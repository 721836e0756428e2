"""Helpers for the `chuckle` CLI tests: the reference's integration scenario
(tests/integration.rs:15-50: DB seeds 100.., query seeds 200..) as BED files."""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "riggle_amd", "chuckle")


def run(*args, check=True):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=600)
    if check and r.returncode != 0:
        raise AssertionError(f"chuckle {' '.join(map(str, args))} failed: {r.stderr}")
    return r


def scenario(orc, tmp, n_db, db_iv, n_q, q_iv, gsize):
    src, qdir = os.path.join(tmp, "src"), os.path.join(tmp, "q")
    os.makedirs(src)
    os.makedirs(qdir)
    for i in range(n_db):
        orc.write_bed(os.path.join(src, f"source_{i}.bed"), *orc.generate_bed(100 + i, db_iv, gsize))
    for i in range(n_q):
        orc.write_bed(os.path.join(qdir, f"query_{i}.bed"), *orc.generate_bed(200 + i, q_iv, gsize))
    return src, qdir


def results(path):
    with open(path) as f:
        return json.load(f)["results"]
