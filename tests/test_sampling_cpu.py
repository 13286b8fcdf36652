"""Scene initialisation sampling (host code in the C-ABI library, no GPU):
Poisson-disk points in a box and the multi-object level-set sampler of
sampling_particles_vdb.hpp:86-160.

The reference draws its box samples with thinks::PoissonDiskSampling
(tph_poisson 0.3), which its build downloads and which is absent here, so the
exact sample positions are parity unpinned.  What a Poisson-disk sampler must
satisfy pins this one instead: no two points closer than r, no probe point of
the box farther than 2r from a sample (Bridson's sampler stops only when
every sample's shell is exhausted), all points inside the box, determinism by
seed.  The level-set sampler is checked against SciPy's trilinear
interpolation (an independent implementation of the reference's BoxSampler,
object_vdb.hpp:73-85), the seed + i convention and the dx rule.
"""
import ctypes

import numpy as np
import pytest
from scipy import ndimage
from scipy.spatial import cKDTree

from pigment_mixing_amd import abi


def test_poisson_box_distance_coverage_and_seed():
    lo, hi, r = np.array([0.0, 0.0, 0.0]), np.array([0.12, 0.05, 0.09]), 0.008
    p = abi.sample_poisson_box(lo, hi, r, 30, 42)
    assert p.shape[0] > 500
    assert np.all(p >= lo.astype(np.float32)) and np.all(p <= hi.astype(np.float32))
    d, _ = cKDTree(p.astype(np.float64)).query(p.astype(np.float64), k=2)
    assert d[:, 1].min() >= r - 1e-7                         # FP32 rounding of the stored points
    probes = np.random.default_rng(1).uniform(lo, hi, (50000, 3))
    dp, _ = cKDTree(p.astype(np.float64)).query(probes)
    assert dp.max() <= 2 * r
    np.testing.assert_array_equal(abi.sample_poisson_box(lo, hi, r, 30, 42), p)
    q = abi.sample_poisson_box(lo, hi, r, 30, 43)
    assert q.shape != p.shape or np.any(q != p)


def sphere_grid(center, radius, voxel, n):
    """Signed distance of a sphere at the voxel points origin + (i, j, k) voxel."""
    origin = np.asarray(center, np.float64) - 0.5 * (n - 1) * voxel
    ax = [origin[a] + voxel * np.arange(n) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    phi = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2) - radius
    return phi.astype(np.float32), origin


def box_grid(half, voxel, n):
    origin = -0.5 * (n - 1) * voxel * np.ones(3)
    ax = [origin[a] + voxel * np.arange(n) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    q = np.stack([np.abs(X) - half[0], np.abs(Y) - half[1], np.abs(Z) - half[2]], -1)
    phi = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)
    return phi.astype(np.float32), origin


def trilinear(phi, voxel, origin, offset, scale, x):
    p = (x.astype(np.float64) - offset) / scale
    idx = ((p - origin) / voxel).T
    inside_grid = np.all((idx >= 0) & (idx <= np.array(phi.shape)[:, None] - 1), axis=0)
    v = ndimage.map_coordinates(phi.astype(np.float64), idx, order=1, mode="nearest")
    return np.where(inside_grid, v, np.inf)


def test_level_sets_follow_the_reference_rules():
    voxel = 0.0025
    s_phi, s_org = sphere_grid((0.0, 0.0, 0.0), 0.02, voxel, 21)
    b_phi, b_org = box_grid((0.015, 0.008, 0.012), voxel, 17)
    objects = [(s_phi, voxel, s_org, np.array([0.05, 0.03, 0.04]), 1.0),
               (b_phi, voxel, b_org, np.array([0.10, 0.02, 0.04]), 1.5)]
    r, seed, ppc = 0.004, 42, 8.0
    x, obj, dx = abi.sample_level_sets(objects, radius=r, attempts=30, seed=seed, ppc=ppc)
    assert set(np.unique(obj)) == {0, 1}
    total_n, total_v = 0, np.float32(0.0)
    f32 = np.float32
    for i, (phi, vox, org, off, sc) in enumerate(objects):
        # the library's FP32 bounds, one rounding per operation
        o, of, s_, v_ = org.astype(f32), off.astype(f32), f32(sc), f32(vox)
        lo = (o * s_ + of).astype(f32)
        hi = ((o + (np.array(phi.shape) - 1).astype(f32) * v_) * s_ + of).astype(f32)
        box = abi.sample_poisson_box(lo, hi, r, 30, seed + i)        # seed + i per object
        total_n += box.shape[0]
        ln = (hi - lo).astype(f32)
        total_v = f32(total_v + f32(f32(ln[0] * ln[1]) * ln[2]))     # float, as the reference
        v = trilinear(phi, vox, org, off, sc, box)
        sure_in, sure_out = v <= -1e-6, v > 1e-6
        mine = x[obj == i]
        # every clearly inside box sample is kept, no clearly outside one is, and
        # the kept ones keep the box sampler's order
        kept = {tuple(p) for p in mine}
        assert all(tuple(p) in kept for p in box[sure_in])
        assert not any(tuple(p) in kept for p in box[sure_out])
        pos = {tuple(p): k for k, p in enumerate(box)}
        idx = np.array([pos[tuple(p)] for p in mine])
        assert np.all(np.diff(idx) > 0)
        assert np.all(trilinear(phi, vox, org, off, sc, mine) <= 1e-6)
    np.testing.assert_allclose(dx, np.cbrt(f32(f32(ppc) / f32(total_n)) * total_v), rtol=3e-7)
    # the sphere's particles lie in the sphere up to the interpolation error
    c = np.array([0.05, 0.03, 0.04])
    assert np.linalg.norm(x[obj == 0] - c, axis=1).max() <= 0.02 + voxel


def test_sampling_errors():
    with pytest.raises(abi.MpmxError) as e:
        abi.sample_poisson_box([0, 0, 0], [0.1, 0.1, 0.1], 0.0)
    assert e.value.code == 1
    with pytest.raises(abi.MpmxError) as e:
        abi.sample_poisson_box([0, 0, 0], [-0.1, 0.1, 0.1], 0.01)
    assert e.value.code == 1
    lib = abi.load()
    lo = (ctypes.c_float * 3)(0, 0, 0)
    hi = (ctypes.c_float * 3)(0.05, 0.05, 0.05)
    n = ctypes.c_int64(0)
    buf = (ctypes.c_float * 30)()
    rc = lib.mpmx_sample_poisson_box(lo, hi, 0.005, 30, 42, buf, 10, ctypes.byref(n))
    assert rc == abi.ERR_CAPACITY and n.value > 10
    phi = np.ones((1, 4, 4), np.float32)
    with pytest.raises(abi.MpmxError) as e:
        abi.sample_level_sets([(phi, 0.01, (0, 0, 0), (0, 0, 0), 1.0)])
    assert e.value.code == 1
