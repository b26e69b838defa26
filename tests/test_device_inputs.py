"""The Philox4x32-10 restatement behind the device input generation, pinned
by the published known-answer vectors of Salmon et al. (SC 2011, the
Random123 distribution's kat_vectors).  CPU only."""
import pytest

from oracle import philox_ref


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, want):
    assert philox_ref.philox4x32_10(ctr, key) == want


def test_stream_uniforms_are_open():
    s = philox_ref.Stream(5, 1, 42)
    u = [s.uniform() for _ in range(64)]
    assert all(0.0 < x < 1.0 for x in u)
