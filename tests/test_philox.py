"""Known-answer vectors of Philox4x32-10 from the Random123 distribution (kat_vectors)."""
import numpy as np

from tests.philox_np import philox4x32_10, philox_normals


def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        got = philox4x32_10(*[np.array([x]) for x in ctr], *key)
        assert tuple(int(x[0]) for x in got) == want


def test_box_muller_statistics():
    n = philox_normals(5, 0, 20, 5000).ravel()
    assert abs(n.mean()) < 0.01 and abs(n.var() - 1) < 0.01
