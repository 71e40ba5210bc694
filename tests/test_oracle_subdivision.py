"""Pins the subdivision oracle (oracle/subdivision.py): dilation against scipy's binary
dilation with a cube, upscaling against the identity and against 2x block repetition,
parameter resampling against exact reproduction of old vertices and of linear and constant
fields, the vertex quantisation against exact arithmetic and index_vertices / index_keys
against hand-computed first-seen indices."""

import numpy as np
import pytest
from scipy import ndimage

from oracle import subdivision as osd
from oracle.extraction import pack_occupancy, unpack_occupancy


def _grid(seed, shape=(6, 7, 64), p=0.05):
    rng = np.random.default_rng(seed)
    return rng.random(shape) < p


def test_write_occupancy_sets_exactly_the_listed_cells():
    rng = np.random.default_rng(0)
    idx = rng.integers(0, 6 * 7 * 64, 300)
    got = unpack_occupancy(osd.write_occupancy(np.concatenate([idx, [-5, 10**9]]), (6, 7, 2)))
    want = np.zeros(6 * 7 * 64, dtype=bool)
    want[idx] = True
    assert np.array_equal(got.reshape(-1), want)


@pytest.mark.parametrize("d", [1, 2, 4, 31])
def test_dilation_is_a_clipped_box(d):
    occ = _grid(1)
    got = unpack_occupancy(osd.dilate_occupancy(pack_occupancy(occ), d))
    want = ndimage.binary_dilation(occ, structure=np.ones((2 * d + 1,) * 3, dtype=bool))
    assert np.array_equal(got, want)


def test_upscale_identity_and_factor_two():
    occ = _grid(2, p=0.2)
    packed = pack_occupancy(occ)
    assert np.array_equal(osd.upscale_occupancy(packed, occ.shape), packed)
    twice = unpack_occupancy(osd.upscale_occupancy(packed, (12, 14, 128)))
    assert np.array_equal(twice, occ.repeat(2, 0).repeat(2, 1).repeat(2, 2))
    # A non-integer ratio covers at least the 2x-style footprint and stays conservative.
    odd = unpack_occupancy(osd.upscale_occupancy(packed, (9, 10, 96)))
    assert odd.sum() >= occ.sum()


def _fields(v, rank, f):
    n = int(v.sum())
    ii, jj, kk = np.nonzero(v)
    I, J, K = (s - 1 for s in v.shape)
    return f(kk / K, jj / J, ii / I).astype(np.float32), n


def test_upscale_parameters_reproduces_old_vertices_and_linear_fields():
    occ = np.ones((4, 5, 32), dtype=bool)
    packed = pack_occupancy(occ)
    v, rank = osd.vertex_occupancy(packed)
    lin = lambda x, y, z: 0.25 + 0.5 * x - 0.75 * y + 0.125 * z  # noqa: E731
    sdf, n = _fields(v, rank, lin)
    sh = np.stack([sdf, 2 * sdf, -sdf], axis=-1)[None]
    new_sdf, new_sh, new_packed = osd.upscale_parameters(sdf, sh, packed, packed, (8, 10, 64))
    nv, _ = osd.vertex_occupancy(new_packed)
    ni, nj, nk = np.nonzero(nv)
    assert np.allclose(new_sdf, lin(nk / 64, nj / 10, ni / 8), atol=2e-6)
    assert np.allclose(new_sh[0, :, 1], 2 * new_sdf, atol=4e-6)
    # New vertices on old vertex positions copy them bit for bit (one weight of 1).
    even = (ni % 2 == 0) & (nj % 2 == 0) & (nk % 2 == 0)
    old = rank[ni[even] // 2, nj[even] // 2, nk[even] // 2]
    assert np.array_equal(new_sdf[even], sdf[old])


def test_upscale_parameters_keeps_constants_under_partial_occupancy():
    occ = _grid(3, shape=(6, 6, 32), p=0.3)
    packed = pack_occupancy(occ)
    v, _ = osd.vertex_occupancy(packed)
    n = int(v.sum())
    sdf = np.full(n, 0.375, dtype=np.float32)
    sh = np.full((4, n, 3), -1.5, dtype=np.float32)
    to = pack_occupancy(_grid(4, shape=(6, 6, 32), p=0.5))
    new_sdf, new_sh, _ = osd.upscale_parameters(sdf, sh, packed, to, (9, 11, 64))
    assert new_sdf.size > 0
    assert np.allclose(new_sdf, 0.375, rtol=1e-6) and np.allclose(new_sh, -1.5, rtol=1e-6)


def test_quantize_is_the_float32_reciprocal_product():
    """oracle.quantize follows torch's device arithmetic, rint(p * fl32(1 / tol)):
    checked against exact arithmetic (a float32 times 100000 is exact in float64, so one
    rounding to float32 gives the float32 product) on points where dividing by tol would
    round the other way."""
    tol = 1e-5
    assert np.float32(1.0 / tol) == np.float32(100000.0)
    pts = np.array([0.260195, 0.470165, 0.706965, 0.25, 0.75, 1e-5, 2.5e-5, -3.5e-5],
                   dtype=np.float32)
    want = [round(float(np.float32(float(p) * 100000.0))) for p in pts]  # ties to even
    assert want[:3] == [26019, 47016, 70696] and want[5:] == [1, 2, -4]
    assert np.array_equal(osd.quantize(pts[:, None], tol)[:, 0], want)
    divided = np.round(pts / np.float32(tol)).astype(np.int64)
    assert not np.array_equal(divided, want)  # the plausible mistake is caught


def test_index_vertices_known_answer():
    """Hand-computed first-seen indices at tol = 0.5 (q = rint(2 p)): ties go to even, -0
    equals 0, and (0, -1, 0) / (3, -1, 2) stay distinct (an unbiased shift-and-or packing of
    negative coordinates would merge them)."""
    pts = np.array([[0, 0, 0], [1, 1, 1], [0.1, 0, 0], [1.2, 0.9, 1.1], [-1, -1, -1],
                    [0.25, 0, 0], [0.75, 0, 0], [-0.0, 0, 0], [0, -0.5, 0], [1.5, -0.5, 1],
                    [-1.1, -0.9, -1.2], [0.0, -0.6, 0.2]], dtype=np.float32)
    want = [0, 1, 0, 1, 4, 0, 6, 0, 8, 9, 4, 8]
    assert osd.index_vertices(pts, 0.5).tolist() == want


def test_index_keys_known_answer():
    keys = np.array([5, -3, 5, 7, -3, 5, -(2**63), 2**63 - 1, -(2**63)], dtype=np.int64)
    assert osd.index_keys(keys).tolist() == [0, 1, 0, 3, 1, 0, 6, 7, 6]
