"""Oracle Counter-V1 RNG against the reference's permanent vectors
(rust/crates/vckss-core/src/rng.rs:366-519)."""
import numpy as np
import pytest

from conftest import load_golden

G = load_golden("rng_vectors.json")
TAGS = {k: int(v, 16) for k, v in G["tags"].items()}


def test_philox_block_permanent_vector(oracle):
    b = G["block"]
    got = oracle.philox_block(b["seed"], TAGS[b["domain"]], b["probe_block"], b["entity"], b["word"])
    assert [f"{x:016x}" for x in got] == b["expected"]


@pytest.mark.parametrize("domain", sorted(G["words"]["expected"]))
def test_domain_word_vectors(oracle, domain):
    w = G["words"]
    got = oracle.rng_word(w["seed"], TAGS[domain], w["probe"], w["entity"], w["word"])
    assert f"{got:016x}" == w["expected"][domain]


def test_compressed_sums(oracle):
    s = G["sums"]
    for trials, expected in s["expected"]:
        assert oracle.rademacher_sum(s["seed"], TAGS[s["domain"]], s["probe"], s["entity"], trials) == expected


def test_batch_split_invariance(oracle):
    # rng.rs:447-476: batches do not change logical atoms
    entity, trials = [3, 11, 29], [1, 20, 65]
    full = oracle.rademacher_fill(987654321, TAGS["target"], 0, 11, entity, trials)
    a = oracle.rademacher_fill(987654321, TAGS["target"], 0, 4, entity, trials)
    b = oracle.rademacher_fill(987654321, TAGS["target"], 4, 7, entity, trials)
    assert np.array_equal(full, np.vstack([a, b]))


def test_packed_fill_matches_scalar(oracle):
    # rng.rs:478-510
    entity, trials = [3, 11, 29, 47], [1, 65, 127, 257]
    for first in range(8):
        for cols in range(1, 9):
            packed = oracle.rademacher_fill(987654321, TAGS["target"], first, cols, entity, trials)
            for c in range(cols):
                for r in range(4):
                    assert packed[c, r] == oracle.rademacher_sum(987654321, TAGS["target"], first + c,
                                                                 entity[r], trials[r])


def test_sum_parity_and_range(oracle):
    # 2*popcount - n has the parity of n and lies in [-n, n]
    for n in [1, 2, 63, 64, 65, 127, 128, 1000]:
        s = oracle.rademacher_sum(5, TAGS["leverage"], 3, 17, n)
        assert (s - n) % 2 == 0 and -n <= s <= n


def test_zero_trials_rejected(oracle):
    with pytest.raises(oracle.OracleError) as e:
        oracle.rademacher_sum(1, TAGS["leverage"], 0, 0, 0)
    assert e.value.code == 70
