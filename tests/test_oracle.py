"""The oracle against the reference's own golden vectors, known-answer tests
and closed forms (CPU only)."""
import numpy as np
import pytest

from tests.cases import flatten, integration_setup, random_files


def test_chacha20_published_vector(orc):
    # ChaCha20, all-zero key and nonce, block 0 (Bernstein's test vector TC1)
    b = orc.chacha_block(np.zeros(8, np.uint32), 0, 0, 20)
    assert [hex(x) for x in b[:4]] == ["0xade0b876", "0x903df1a0", "0xe56a5d40", "0x28bd8653"]
    assert hex(b[15]) == "0x8665eeb2"


def test_golden_intervals_hits_and_bp(orc):
    # tests/integration.rs:57-75: 158 hits, 413,652 bp
    db, q = integration_setup(orc, 10, 2000, 1, 1000, 50_000_000)
    c, bp = orc.sweep_bruteforce(db, q, 10, 1)
    assert c.sum() == 158 and bp.sum() == 413_652


def test_golden_counts_nonzero_pairs(orc):
    # tests/integration.rs:103-115: nnz 10, total 58
    db, q = integration_setup(orc, 5, 1000, 2, 500, 30_000_000)
    c, _ = orc.sweep_bruteforce(db, q, 5, 2)
    assert (c > 0).sum() == 10 and c.sum() == 58


def test_layered_reference_path_matches_exact(orc):
    db, q = integration_setup(orc, 5, 1000, 2, 500, 30_000_000)
    L = orc.LayeredDB(*db, 24, 5)
    cnt, bins, _ = L.query(*q, 2)
    c, bp = orc.sweep_bruteforce(db, q, 5, 2)
    assert (cnt == c).all()
    assert np.allclose(bins, bp / 100.0, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("seed", [0, 1])
def test_rank_identity_matches_bruteforce(orc, seed):
    rng = np.random.default_rng(seed)
    db = flatten(random_files(rng, 6, 3, (0, 300), 10_000, 1_500))
    q = flatten(random_files(rng, 3, 3, (0, 200), 10_000, 2_000))
    a = orc.sweep_bruteforce(db, q, 6, 3)
    b = orc.sweep_ranks(db, q, 6, 3, 3)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


def test_layer_config(orc):
    # src/io/meta.rs:152-170
    assert [orc.layer_for_size(s) for s in (0, 127, 128, 511, 512, 2047, 2048)] == [0, 0, 1, 1, 2, 2, 3]
    assert [orc.layer_max_size(k) for k in (0, 1, 2)] == [128, 512, 2048]


def test_jump_table(orc):
    # src/io/jump.rs:181-240
    assert orc.build_jump_table([], 64).tolist() == [0]
    assert orc.build_jump_table([10, 50, 80], 100).tolist() == [0]
    assert orc.build_jump_table([10, 150, 250], 100).tolist() == [0, 1, 2]
    assert orc.build_jump_table([10, 350], 100).tolist() == [0, 1, 1, 1]
