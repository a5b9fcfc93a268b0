"""Pin the oracle's Gaussian generator (src_random/RNGs.c restated).

Checks: the published MT19937 check value, bit-for-bit agreement with the
reference's own RNGs.c compiled from /root/reference into oracle/_ref (when
it was built), Acklam's published error bound against scipy's ndtri, and the
numpy generator of lowmachmixtures_amd.problems against the oracle stream.
"""
import os

import numpy as np
from scipy import stats
import pytest

from lowmachmixtures_amd import problems as P
from oracle import oracle as O


def test_mt19937_published_check_value():
    # init_genrand(5489): the 10000th output is 4123659995 (MT19937 reference value)
    lib = O.lib()
    st = (O.C.c_uint32 * 625)()
    lib.orc_mt_seed(st, O.C.c_uint32(5489))
    v = 0
    for _ in range(10000):
        v = lib.orc_mt_next(st)
    assert v == 4123659995


def test_mt19937_first_outputs_seed_5489():
    # first output of init_genrand(5489) is 3499211612
    raw = P.mt_raw(5489, 3)
    assert int(raw[0]) == 3499211612


@pytest.mark.parametrize("seed", [1, 5489, 12345, 2 ** 31 + 7])
def test_oracle_mt_matches_reference_rngs_c(seed):
    ref = O.ref_rng()
    if ref is None:
        pytest.skip("reference RNGs.c not built into oracle/_ref (reference tree absent)")
    ref.srandgen(seed)
    want = [ref.genrandbase() & 0xFFFFFFFF for _ in range(2000)]
    got = P.mt_raw(seed, 2000)
    assert [int(x) for x in got] == want


def test_oracle_normals_match_reference_genrandn():
    ref = O.ref_rng()
    if ref is None:
        pytest.skip("reference RNGs.c not built into oracle/_ref")
    ref.srandgen(77)
    out = O.C.c_double()
    want = []
    for _ in range(5000):
        ref.genrandn(O.C.byref(out))
        want.append(out.value)
    got = O.fill_normal(77, 5000)
    assert np.array_equal(got, np.array(want))


def test_acklam_error_bound():
    from scipy.special import ndtri

    p = np.concatenate([np.linspace(1e-9, 0.02425, 2000), np.linspace(0.02425, 0.97575, 4000),
                        np.linspace(0.97575, 1 - 1e-9, 2000)])
    x = np.array([O.inv_norm_dist(float(v)) for v in p])
    ref = ndtri(p)
    mask = np.abs(ref) > 1e-3
    rel = np.abs(x[mask] - ref[mask]) / np.abs(ref[mask])
    assert rel.max() < 1.15e-9
    assert np.all(np.abs(x[~mask] - ref[~mask]) < 1.2e-12)


def test_numpy_generator_matches_oracle_stream():
    a = P.normals(4242, 20000)
    b = O.fill_normal(4242, 20000)
    # central region uses only + - * /: identical; tails use log/sqrt of the platform
    assert np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))) < 1e-15
    assert np.mean(a == b) > 0.99


def test_inverse_normal_symmetry():
    # 1 - p is rounded, so the two tails agree to Acklam's bound, not to ulps
    for p in (1e-6, 0.01, 0.2, 0.4999):
        assert O.inv_norm_dist(p) == pytest.approx(-O.inv_norm_dist(1 - p), rel=2.5e-9)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rngs_c_vectors.npz")


@pytest.mark.parametrize("seed", [1, 5489, 12345, 2 ** 31 + 7])
def test_oracle_mt_matches_committed_reference_vectors(seed):
    g = np.load(GOLDEN)
    assert np.array_equal(P.mt_raw(seed, 2000), g[f"raw_{seed}"])


def test_oracle_normals_match_committed_reference_vectors():
    g = np.load(GOLDEN)
    assert np.array_equal(O.fill_normal(77, 5000), g["normal_77"])


# ---------------- counter-based stream (Philox4x32-10 + Box-Muller) ----------------

# Known-answer vectors of Philox4x32-10 published with the Random123 library
# (Salmon et al., SC'11): (counter, key) -> output words.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(O.philox4x32_10(ctr, key)) == want


def test_philox_stream_block_to_normal_pair():
    # value 2b, 2b+1 of (seed, step) come from the block of counter (b, step)
    # under the key seed: words (x1:x0) and (x3:x2) -> 53-bit uniforms
    seed, step, b = 0x1234_5678_9ABC_DEF0, 7, 5
    x = O.philox4x32_10((b, 0, step, 0), (seed & 0xFFFFFFFF, seed >> 32))
    u1 = (float(((x[1] << 32) | x[0]) >> 11) + 0.5) / 2.0 ** 53
    u2 = (float(((x[3] << 32) | x[2]) >> 11) + 0.5) / 2.0 ** 53
    r = np.sqrt(-2.0 * np.log(u1))
    got = O.fill_normal_philox(seed, step, 2 * b, 2)
    np.testing.assert_allclose(got, [r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)], rtol=1e-14)


def test_philox_stream_offsets_and_independence():
    full = O.fill_normal_philox(11, 3, 0, 4001)
    for v0 in (1, 2, 999):  # any start, odd or even, gives the same values
        assert np.array_equal(O.fill_normal_philox(11, 3, v0, 4001 - v0), full[v0:])
    other_step = O.fill_normal_philox(11, 4, 0, 4001)
    other_seed = O.fill_normal_philox(12, 3, 0, 4001)
    assert abs(np.corrcoef(full, other_step)[0, 1]) < 0.1
    assert abs(np.corrcoef(full, other_seed)[0, 1]) < 0.1


def test_philox_stream_is_standard_normal():
    z = O.fill_normal_philox(2024, 0, 0, 1 << 20)
    n = z.size
    assert abs(z.mean()) < 5.0 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
    assert abs(np.corrcoef(z[:-1], z[1:])[0, 1]) < 5.0 / np.sqrt(n)  # neighbours, incl. Box-Muller pairs
    assert stats.kstest(z, "norm").pvalue > 1e-3
