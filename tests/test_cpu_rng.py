"""The oracle's counter-based random streams, pinned by published known-answer
vectors and by properties of the derived variates.

Per photon and stage the stream is Philox4x32-10 keyed by the 64-bit seed,
counter (tag, stage, block, source id); each block yields four 32-bit words,
used in order, each as u32 / (2^32 - 1).  Gaussians are Marsaglia's polar
method with the second variate of each accepted pair cached, the scheme of
jdmath/src/gaussrnd.c:30-67 (JDMgaussian_random)."""
import numpy as np
import pytest

from oracle import oracle as O

# Random123 (Salmon, Moraes, Dror, Shaw, SC'11) known-answer vectors for
# Philox4x32-10: (counter, key) -> output.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(int(v) for v in O.philox(ctr, key)) == expect


def test_uniform_stream_layout():
    """Uniform k of a stream is word k % 4 of block k // 4, scaled by 1/(2^32-1);
    the counter is (tag, stage, block, source id) and the key the seed."""
    seed, tag, stage, src = 0x0123456789ABCDEF, 77, 3, 5
    u = O.stream_uniforms(seed, tag, stage, 12, source_id=src)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for blk in range(3):
        words = O.philox((tag, stage, blk, src), key)
        np.testing.assert_array_equal(u[4 * blk:4 * blk + 4], words.astype(np.float64) / 4294967295.0)


def test_streams_are_distinct_per_tag_stage_source():
    base = O.stream_uniforms(9, 1, 3, 8)
    for args in ((9, 2, 3), (9, 1, 4), (10, 1, 3)):
        assert not np.array_equal(base, O.stream_uniforms(*args, 8))
    assert not np.array_equal(base, O.stream_uniforms(9, 1, 3, 8, source_id=1))


def test_uniform_moments():
    u = np.concatenate([O.stream_uniforms(1, t, 3, 64) for t in range(2000)])
    assert u.min() >= 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.005
    assert abs(u.var() - 1.0 / 12.0) < 0.002


def test_gaussian_pair_is_polar_method():
    """The first two Gaussians of a stream come from the first accepted
    polar pair: g0 / g1 = v0 / v1 with v = 2u - 1 and v0^2 + v1^2 in (0, 1),
    and g0^2 + g1^2 = -2 ln(s) for s = v0^2 + v1^2."""
    for tag in range(200):
        u = O.stream_uniforms(5, tag, 3, 64)
        g = O.stream_gaussians(5, tag, 3, 2)
        v = 2.0 * u - 1.0
        k = 0
        while True:
            s = v[k] ** 2 + v[k + 1] ** 2
            if 0.0 < s < 1.0:
                break
            k += 2
        assert g[0] * v[k + 1] == pytest.approx(g[1] * v[k], rel=1e-12, abs=1e-15)
        assert g[0] ** 2 + g[1] ** 2 == pytest.approx(-2.0 * np.log(s), rel=1e-12)


def test_gaussian_moments():
    g = np.concatenate([O.stream_gaussians(3, t, 5, 50) for t in range(4000)])
    assert abs(g.mean()) < 0.01
    assert abs(g.std() - 1.0) < 0.01
    assert abs(np.mean(g ** 4) - 3.0) < 0.1
