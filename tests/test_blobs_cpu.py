"""Scene initialisation of the paint app (host code, CPU): the spatula's
animation file (ObjectSpatula::loadAnimation, object_spatula.hpp:235-250) and
its keyframe poses (lerp / slerp, :164-233), and the blob scene
(Simulation::blobs, simulation_blobs.cpp:7-218) on a database of concentric
spheres, where the morph, the volume bisection and the placement have closed
forms."""
import json
import math
import os

import numpy as np
import pytest

from pigment_mixing_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
HEAD = os.path.join(HERE, "golden", "spatula_motion_head.bin")
FULL = "/root/reference/matter/animations/spatula_motion.bin"


# ------------------------------------------------------------------ animation file
def test_keyframe_file_loads_as_framedata_records():
    fr = abi.load_keyframes(HEAD)
    raw = np.frombuffer(open(HEAD, "rb").read(), "<f4").reshape(-1, 7)     # {pos[3], quat xyzw[4]}
    assert fr.shape == (181, 7)
    np.testing.assert_array_equal(fr, raw)
    np.testing.assert_allclose(np.linalg.norm(fr[:, 3:], axis=1), 1.0, atol=1e-6)
    lib = abi.load()
    n = abi.ctypes.c_int64(0)
    small = np.zeros((10, 7), np.float32)
    assert lib.mpmx_load_keyframes(os.fsencode(HEAD), abi.fptr(small), 10, abi.ctypes.byref(n)) == abi.ERR_CAPACITY
    assert n.value == 181 and not small.any()
    assert lib.mpmx_load_keyframes(b"/nonexistent/anim.bin", None, 0, abi.ctypes.byref(n)) == 6   # MPMX_ERR_IO


@pytest.mark.skipif(not os.path.exists(FULL), reason="the reference's animation file is not mounted")
def test_reference_animation_file_record_count():
    meta = json.load(open(os.path.join(HERE, "golden", "spatula_motion.json")))
    fr = abi.load_keyframes(FULL)
    assert fr.shape[0] == meta["records"] == 3601
    np.testing.assert_array_equal(fr[:181], abi.load_keyframes(HEAD))


def _slerp(q0, q1, t):
    """Eigen's Quaternion::slerp (shortest arc), quaternions as (x, y, z, w)."""
    d = float(np.dot(q0, q1))
    ad = abs(d)
    if ad >= 1.0 - 1e-12:
        s0, s1 = 1.0 - t, t
    else:
        th = math.acos(ad)
        s = math.sin(th)
        s0, s1 = math.sin((1 - t) * th) / s, math.sin(t * th) / s
    if d < 0:
        s1 = -s1
    return s0 * q0 + s1 * q1


def _rot(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_keyframe_poses_lerp_and_slerp_of_the_file():
    """During step k the knife is at time k dt of the animation played at 60 fps:
    position lerp and rotation slerp between the bracketing records, at and
    between record boundaries (ObjectSpatula::move, object_spatula.hpp:164-233);
    its velocity is the backward difference of positions over dt."""
    fr = abi.load_keyframes(HEAD).astype(np.float64)
    lib = abi.load()
    k = abi.Knife()
    k.shape, k.motion = abi.KNIFE_SPATULA, abi.MOTION_KEYFRAMES
    k.half[:] = [0.368, 0.07, 0.736]
    k.offset = 0.05
    dt = 1.0 / 1200.0                 # 20 steps per record
    f32 = fr.astype(np.float32)
    moving = np.where(np.abs(np.diff(fr[:, :3], axis=0)).max(1) > 1e-6)[0]
    assert len(moving) > 20
    steps = []
    for rec in moving[:: max(1, len(moving) // 6)]:
        steps += [20 * int(rec), 20 * int(rec) + 7, 20 * int(rec) + 10]
    for step in steps:
        st = abi.KnifeState()
        assert lib.mpmx_knife_state_eval(abi.ctypes.byref(k), abi.fptr(f32), f32.shape[0], 60.0, dt, step,
                                         abi.ctypes.byref(st)) == 0
        t = step * dt * 60.0
        i0 = int(math.floor(t))
        i1 = min(i0 + 1, fr.shape[0] - 1)
        a = t - i0
        p = fr[i0, :3] + a * (fr[i1, :3] - fr[i0, :3])
        R = _rot(_slerp(fr[i0, 3:], fr[i1, 3:], a))
        T = np.array(st.T).reshape(3, 4)
        np.testing.assert_allclose(T[:, 3], p, atol=2e-6)
        np.testing.assert_allclose(T[:, :3], R, atol=2e-6)
        Ti = np.array(st.Tinv).reshape(3, 4)
        np.testing.assert_allclose(Ti[:, :3] @ T[:, :3], np.eye(3), atol=2e-6)
        np.testing.assert_allclose(Ti[:, :3] @ p + Ti[:, 3], 0.0, atol=2e-6)
        tp = (step - 1) * dt * 60.0
        j0 = int(math.floor(tp))
        pp = fr[j0, :3] + (tp - j0) * (fr[min(j0 + 1, fr.shape[0] - 1), :3] - fr[j0, :3])
        np.testing.assert_allclose(np.array(st.v), (p - pp) / dt, rtol=2e-3, atol=2e-3)


# ------------------------------------------------------------------ blob scene
VOXEL = 0.03
VOLUMES = [0.0056, 0.0151, 0.0246, 0.0341, 0.0436, 0.053]      # the reference's database (simulation.hpp:272-279)
CENTER = np.array([0.0, 0.3, 0.0])


def _sphere_db():
    """Concentric spheres of the database's volumes on one grid (voxel 0.03);
    calibration = volume / voxel-counted volume, as a calibration factor is."""
    r_max = (3 * VOLUMES[-1] / (4 * math.pi)) ** (1 / 3)
    m = int(math.ceil((r_max + 4 * VOXEL) / VOXEL))
    idx = np.arange(-m, m + 1)
    org = CENTER / VOXEL
    org = np.round(org) * VOXEL - m * VOXEL
    g = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1) * VOXEL + (org + m * VOXEL)
    dist = np.linalg.norm(g - CENTER, axis=-1)
    db = []
    for V in VOLUMES:
        r = (3 * V / (4 * math.pi)) ** (1 / 3)
        phi = (dist - r).astype(np.float32)
        count = int((phi <= 0).sum())
        db.append((phi, VOXEL, tuple(org), V, V / (count * VOXEL ** 3)))
    return db, dist


def _voxel_volume(dist, rA, rB, fA, fB, t):
    phi = (1 - t) * (dist - rA) + t * (dist - rB)
    return float((phi <= 0).sum()) * VOXEL ** 3 * ((1 - t) * fA + t * fB)


def test_blob_volume_morph_and_bisection():
    db, dist = _sphere_db()
    ratios = [0.05, 0.3, 0.55, 0.95]
    x, obj, dx, place = abi.blob_scene(db, ratios[:1])    # one blob: scene and placement run
    assert x.shape[0] > 1000 and dx > 0
    for r in ratios:
        _, _, _, (p,) = abi.blob_scene(db, [r], radius=0.03)
        rr = min(0.9, max(0.1, r))
        target = VOLUMES[0] + (rr - 0.1) / 0.8 * (VOLUMES[-1] - VOLUMES[0])
        assert abs(p.target_volume - target) <= 1e-6 * target
        pair = next(i for i in range(5) if VOLUMES[i] <= target <= VOLUMES[i + 1])
        assert p.pair == pair
        rA, rB = [(3 * VOLUMES[i] / (4 * math.pi)) ** (1 / 3) for i in (pair, pair + 1)]
        fA, fB = db[pair][4], db[pair + 1][4]
        # the blend of concentric spheres is the sphere of radius (1 - t) rA + t rB:
        # its voxel-counted, calibrated volume meets the target to the jump of
        # the count over the bisection's last bracket
        vol = _voxel_volume(dist, rA, rB, fA, fB, p.t)
        lo = _voxel_volume(dist, rA, rB, fA, fB, max(0.0, p.t - 2.0 ** -14))
        hi = _voxel_volume(dist, rA, rB, fA, fB, min(1.0, p.t + 2.0 ** -14))
        assert lo - 1e-9 <= target <= hi + 1e-9 or abs(vol - target) <= 1e-6 * target, (r, vol, target, lo, hi)
        # and t is that of the closed-form radius up to the voxel discretisation
        r_t = (1 - p.t) * rA + p.t * rB
        r_exact = (3 * target / (4 * math.pi)) ** (1 / 3)
        assert abs(r_t - r_exact) <= 0.5 * VOXEL


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_blob_placement_touching_and_centred(n):
    """Blobs touch pairwise where the reference's closed forms place them: n=2
    side by side, n=3 a triangle of tangent discs, n=4 two tangent pairs at a
    common z, n>4 a row; the x-z bounds of the discs are centred; particles
    lie in their own blob."""
    db, _ = _sphere_db()
    ratios = [0.2, 0.8, 0.5, 0.35, 0.65][:n]
    x, obj, dx, place = abi.blob_scene(db, ratios, radius=0.02)
    c = np.array([[0.5 * (p.bbox_min[0] + p.bbox_max[0]), 0.5 * (p.bbox_min[2] + p.bbox_max[2])] for p in place])
    r = np.array([p.radius for p in place])
    d = lambda i, j: float(np.linalg.norm(c[i] - c[j]))
    tol = 1e-5
    if n == 2:
        assert abs(d(0, 1) - (r[0] + r[1])) < tol
    elif n == 3:
        for i, j in ((0, 1), (0, 2), (1, 2)):
            assert abs(d(i, j) - (r[i] + r[j])) < tol
    elif n == 4:
        assert abs(d(0, 1) - (r[0] + r[1])) < tol and abs(d(2, 3) - (r[2] + r[3])) < tol
        assert abs(c[2, 1] - c[3, 1]) < tol and abs(c[0, 1] - c[1, 1]) < tol
        # the rows sit as close as the tangency of the nearest cross pair allows
        cross = min(d(i, j) - (r[i] + r[j]) for i in (0, 1) for j in (2, 3))
        assert abs(cross) < tol
    else:
        for i in range(n - 1):
            assert abs(d(i, i + 1) - (r[i] + r[i + 1])) < tol
    lo = np.min(c - r[:, None], axis=0)
    hi = np.max(c + r[:, None], axis=0)
    np.testing.assert_allclose(0.5 * (lo + hi), 0.0, atol=1e-5)
    # each particle inside its own (morphed) sphere: centre = placed bbox centre
    for i, p in enumerate(place):
        ctr = 0.5 * (np.array(p.bbox_min) + np.array(p.bbox_max))
        rA, rB = [(3 * VOLUMES[k] / (4 * math.pi)) ** (1 / 3) for k in (p.pair, p.pair + 1)]
        r_t = (1 - p.t) * rA + p.t * rB
        mine = x[obj == i]
        assert mine.shape[0] > 100
        assert np.linalg.norm(mine - ctr, axis=1).max() <= r_t + VOXEL
