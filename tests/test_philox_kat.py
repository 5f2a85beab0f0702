"""Known-answer tests of the device RNG's Philox4x32-10: the oracle
restatement and the library's own block function (evaluated on the host,
same source as the device code) against the published Random123 vectors
for philox4x32 with 10 rounds."""
import numpy as np
import pytest

from oracle import philox as PX

# counter (4 words), key (2 words) -> output (4 words)
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_oracle_matches_published_vectors(ctr, key, want):
    assert PX.philox4x32_10(ctr, key) == want


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_library_block_matches_published_vectors(ctr, key, want):
    from hmclab_amd.engine import philox4x32_10

    assert philox4x32_10(ctr, key) == want


def test_library_block_matches_oracle_on_random_counters():
    from hmclab_amd.engine import philox4x32_10

    rng = np.random.default_rng(0)
    for _ in range(200):
        ctr = [int(v) for v in rng.integers(0, 2 ** 32, size=4)]
        key = [int(v) for v in rng.integers(0, 2 ** 32, size=2)]
        assert philox4x32_10(ctr, key) == PX.philox4x32_10(ctr, key)


def test_conversions():
    # u53 covers [0, 1) with 2^-53 resolution; the Box-Muller u1 never reaches 0
    assert PX.u53(0, 0) == 0.0
    assert PX.u53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    z = np.array([PX.normal_pair(3, c, 7, q) for c in range(64) for q in range(64)]).ravel()
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1.0) < 0.05
