"""Philox4x32-10 and Philox2x32-10 of the oracle against the published Random123
known-answer vectors (Salmon et al. SC'11; kat_vectors of the Random123
distribution)."""
import numpy as np

import orc

KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


# philox2x32 10: counter, key, output
KAT2 = [
    ((0x00000000, 0x00000000), 0x00000000, (0xFF1DAE59, 0x6CD10DF2)),
    ((0xFFFFFFFF, 0xFFFFFFFF), 0xFFFFFFFF, (0x2C3F628B, 0xAB4FD7AD)),
    ((0x243F6A88, 0x85A308D3), 0x13198A2E, (0xDD7CE038, 0xF62A4C12)),
]


def test_known_answer_vectors():
    for ctr, key, want in KAT:
        assert tuple(orc.philox(ctr, key)) == want


def test_known_answer_vectors_2x32():
    for ctr, key, want in KAT2:
        assert orc.philox2(ctr, key) == want


def test_uniform_open_interval_and_moments():
    lib = orc.lib()
    assert 0.0 < lib.orc_u01(0) < 1e-6
    assert 1.0 - 1e-6 < lib.orc_u01(0xFFFFFFFF) < 1.0
    u = np.array([lib.orc_u01(x) for w in range(2000) for x in orc.philox((w, 7, 0, 0), (1, 2))])
    assert abs(u.mean() - 0.5) < 0.02
    assert abs(u.var() - 1.0 / 12.0) < 0.01
