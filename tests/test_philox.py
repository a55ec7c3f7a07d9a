"""Counter-based noise (Philox4x32-10 + Box-Muller), host side of the C ABI.

* Philox4x32-10 against the known-answer vectors published with the Random123 library
  (Salmon, Moraes, Dror, Shaw, SC'11; kat_vectors, philox4x32 with 10 rounds): the
  all-zero, all-ones and pi-digit counter/key pairs.
* The normals: their uniforms follow from the block by exact integer arithmetic, checked
  here against a numpy restatement of the 53-bit construction; moments over 2^20 draws.
"""
import numpy as np
import pytest

from spsslangevinintegration_amd import _capi

KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    got = _capi.philox4x32_10(ctr, key)
    assert [int(v) for v in got] == list(want)


def test_philox_normals_construction():
    seed, step, block = 0x5EED0003, 12345678901, 1
    n = 64
    z0, z1 = _capi.philox_normals(seed, step, n, block=block, first=1000)
    for k in range(n):
        r = _capi.philox4x32_10((1000 + k, block, step & 0xffffffff, step >> 32), (seed & 0xffffffff, seed >> 32))
        r = [int(v) for v in r]
        a = ((r[0] << 21) ^ (r[1] >> 11)) & ((1 << 53) - 1)
        b = ((r[2] << 21) ^ (r[3] >> 11)) & ((1 << 53) - 1)
        u1 = (a + 1.0) / 2.0 ** 53
        u2 = b / 2.0 ** 53
        rad = np.sqrt(-2.0 * np.log(u1))
        assert abs(z0[k] - rad * np.cos(2 * np.pi * u2)) <= 1e-14 * max(1.0, rad)
        assert abs(z1[k] - rad * np.sin(2 * np.pi * u2)) <= 1e-14 * max(1.0, rad)


def test_philox_normals_moments():
    n = 1 << 20
    z0, z1 = _capi.philox_normals(7, 3, n)
    for z in (z0, z1):
        assert abs(z.mean()) < 5.0 / np.sqrt(n)
        assert abs(z.var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
        assert abs(np.mean(z ** 4) - 3.0) < 5.0 * np.sqrt(96.0 / n)
    assert abs(np.mean(z0 * z1)) < 5.0 / np.sqrt(n)
    # different steps, blocks and particles give different draws
    a, _ = _capi.philox_normals(7, 4, 16)
    b, _ = _capi.philox_normals(7, 3, 16, block=1)
    assert not np.allclose(a, z0[:16]) and not np.allclose(b, z0[:16])
