"""The numpy restatement of Philox4x32-10 used to check device-generated
streams, against the published known-answer vectors of the Random123
library (kat_vectors, philox4x32 with 10 rounds)."""

import numpy as np

from philox_ref import philox4x32_10


def _run(ctr, key):
    out = philox4x32_10(*[[c] for c in ctr], *key)
    return [int(x[0]) for x in out]


def test_philox4x32_10_known_answers():
    assert _run([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert _run([0xFFFFFFFF] * 4, [0xFFFFFFFF, 0xFFFFFFFF]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert _run([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_vectorised_equals_scalar():
    rng = np.random.default_rng(0)
    c = rng.integers(0, 1 << 32, (4, 8), dtype=np.uint64)
    out = philox4x32_10(*c, 123, 456)
    for i in range(8):
        one = philox4x32_10(*[[x[i]] for x in c], 123, 456)
        assert [int(o[0]) for o in one] == [int(o[i]) for o in out]
