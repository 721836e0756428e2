"""Seeded input builders shared by the tests."""
import numpy as np


def flatten(files):
    """files: list of (chrom/shard uint array, start, end) -> (shard, start, end, sid)."""
    if not files:
        z = np.zeros(0, np.uint32)
        return z, z, z, z
    sh = np.concatenate([np.asarray(f[0], np.uint32) for f in files])
    s = np.concatenate([np.asarray(f[1], np.uint32) for f in files])
    e = np.concatenate([np.asarray(f[2], np.uint32) for f in files])
    sid = np.concatenate([np.full(len(f[1]), i, np.uint32) for i, f in enumerate(files)])
    return sh, s, e, sid


def integration_setup(orc, n_db, db_iv, n_q, q_iv, gsize):
    """tests/integration.rs:15-50 of the reference: DB seeds 100.., query seeds 200.."""
    db = [orc.generate_bed(100 + i, db_iv, gsize) for i in range(n_db)]
    q = [orc.generate_bed(200 + i, q_iv, gsize) for i in range(n_q)]
    return flatten(db), flatten(q)


def random_files(rng, nfiles, nshards, per_file, span, max_len, min_len=1):
    files = []
    for _ in range(nfiles):
        n = int(per_file) if np.isscalar(per_file) else int(rng.integers(*per_file))
        sh = rng.integers(0, nshards, n).astype(np.uint32)
        ln = rng.integers(min_len, max_len + 1, n).astype(np.uint32)
        s = rng.integers(0, span, n).astype(np.uint32)
        files.append((sh, s, s + ln))
    return files
