"""ctypes binding of include/mpmx.h (libmpmx.so).

The library is the product: every call here goes through the C-ABI into the
HIP path.  `load()` raises if libmpmx.so is missing — there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmpmx.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_i64_p = ctypes.POINTER(ctypes.c_int64)

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_DEVICE, ERR_CAPACITY, ERR_IO = range(7)
BC_NOSLIP, BC_SLIPSTICK, BC_SLIPFREE = 0, 1, 2
PLATE_TOP, PLATE_BOTTOM, PLATE_LEFT, PLATE_RIGHT, PLATE_FRONT, PLATE_BACK = range(6)
PLASTIC_NONE, PLASTIC_VMVISC, PLASTIC_DPVISC = 0, 5, 6
ERR_CAPACITY = 5   # MPMX_ERR_CAPACITY
CLOSURE_EXACT, CLOSURE_LEARNED = 0, 1
KNIFE_SLAB, KNIFE_SPATULA = 0, 1
MOTION_STATIC, MOTION_LINEAR, MOTION_CIRCULAR, MOTION_LINEAR_THEN_CIRCULAR, MOTION_KEYFRAMES = range(5)


class LevelSet(ctypes.Structure):
    """mpmx_level_set: a dense signed-distance grid placed as p * scale + offset."""
    _fields_ = [("phi", ctypes.POINTER(ctypes.c_float)), ("dims", ctypes.c_int32 * 3), ("voxel", ctypes.c_float),
                ("origin", ctypes.c_float * 3), ("offset", ctypes.c_float * 3), ("scale", ctypes.c_float)]


class BlobModel(ctypes.Structure):
    """mpmx_blob_model: one model of the blob database."""
    _fields_ = [("shape", LevelSet), ("volume", ctypes.c_float), ("calibration", ctypes.c_float)]


class BlobPlacement(ctypes.Structure):
    _fields_ = [("pair", ctypes.c_int32), ("t", ctypes.c_float), ("target_volume", ctypes.c_float),
                ("radius", ctypes.c_float), ("offset", ctypes.c_float * 3), ("bbox_min", ctypes.c_float * 3),
                ("bbox_max", ctypes.c_float * 3)]


class Config(ctypes.Structure):
    _fields_ = [
        ("E", ctypes.c_float), ("nu", ctypes.c_float), ("rho", ctypes.c_float),
        ("flip_ratio", ctypes.c_float),
        ("plastic_model", ctypes.c_int),
        ("q_max", ctypes.c_float), ("q_min", ctypes.c_float),
        ("xi", ctypes.c_float), ("p_min", ctypes.c_float),
        ("perzyna_exp", ctypes.c_float), ("perzyna_visc", ctypes.c_float),
        ("q_prefac", ctypes.c_float),
        ("cfl", ctypes.c_float), ("M", ctypes.c_float), ("q_cohesion", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("D_max", ctypes.c_float), ("D_edge0", ctypes.c_float), ("D_edge1", ctypes.c_float),
        ("start_boost_time", ctypes.c_float), ("end_boost_time", ctypes.c_float),
        ("boost_factor", ctypes.c_float),
        ("dx", ctypes.c_float), ("ppc", ctypes.c_float),
        ("grid_dims", ctypes.c_int * 3),
        ("origin_mode", ctypes.c_int),
        ("grid_origin", ctypes.c_float * 3),
        ("grid_reference_point", ctypes.c_float * 3),
        ("dt", ctypes.c_float), ("fps", ctypes.c_float),
        ("use_musl", ctypes.c_int),
        ("use_particle_boundaries", ctypes.c_int),
        ("particle_boundary_min", ctypes.c_float * 3),
        ("particle_boundary_max", ctypes.c_float * 3),
        ("closure", ctypes.c_int), ("closure_tol", ctypes.c_float),
        ("n_threads", ctypes.c_int),
        ("use_pradhana", ctypes.c_int),
    ]


class Plate(ctypes.Structure):
    _fields_ = [
        ("pos_object", ctypes.c_float), ("plate_type", ctypes.c_int), ("bc", ctypes.c_int),
        ("friction", ctypes.c_float), ("pos_lower", ctypes.c_float), ("pos_upper", ctypes.c_float),
        ("v", ctypes.c_float * 3), ("vmin_factor", ctypes.c_float), ("load_factor", ctypes.c_float),
    ]


class Knife(ctypes.Structure):
    _fields_ = [
        ("shape", ctypes.c_int), ("half", ctypes.c_float * 3), ("offset", ctypes.c_float),
        ("bc", ctypes.c_int), ("friction", ctypes.c_float), ("motion", ctypes.c_int),
        ("p0", ctypes.c_float * 3), ("tilt_axis", ctypes.c_float * 3), ("tilt_angle", ctypes.c_float),
        ("velocity", ctypes.c_float * 3), ("radius", ctypes.c_float), ("speed", ctypes.c_float),
        ("phase0", ctypes.c_float), ("t_switch", ctypes.c_float),
    ]


class KnifeState(ctypes.Structure):
    _fields_ = [
        ("T", ctypes.c_float * 12), ("Tinv", ctypes.c_float * 12), ("Nmat", ctypes.c_float * 9),
        ("v", ctypes.c_float * 3), ("w", ctypes.c_float * 3),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("step", ctypes.c_int64), ("time", ctypes.c_double), ("n_particles", ctypes.c_int64),
        ("n_active", ctypes.c_int64), ("plastic_count", ctypes.c_int64),
        ("fallback_count", ctypes.c_int64), ("active_nodes", ctypes.c_int64),
        ("active_blocks", ctypes.c_int64), ("stray_count", ctypes.c_int64),
        ("device_error", ctypes.c_uint32), ("ms_last_step", ctypes.c_float),
        ("learned_count", ctypes.c_int64), ("newton_iters", ctypes.c_int64),
    ]


class KernelTimes(ctypes.Structure):
    _fields_ = [
        ("ms_g2p", ctypes.c_float), ("ms_sort", ctypes.c_float), ("ms_p2g", ctypes.c_float),
        ("ms_grid", ctypes.c_float), ("ms_step", ctypes.c_float), ("steps", ctypes.c_int64),
        ("particles_per_step", ctypes.c_double), ("active_nodes_per_step", ctypes.c_double),
        ("ms_closure", ctypes.c_float), ("queued_per_step", ctypes.c_double),
    ]


class Box(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_float * 3), ("cells", ctypes.c_int * 3), ("latent", ctypes.c_float * 7)]


class TuneReport(ctypes.Structure):
    _fields_ = [("wg_p2g", ctypes.c_int), ("wg_g2p", ctypes.c_int), ("wg_grid", ctypes.c_int),
                ("ms_p2g", ctypes.c_float), ("ms_g2p", ctypes.c_float), ("ms_grid", ctypes.c_float),
                ("trials", ctypes.c_int), ("steps", ctypes.c_int64)]


class TrainConfig(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("batch", ctypes.c_int), ("lr", ctypes.c_float),
                ("lr_final", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float), ("seed", ctypes.c_uint64), ("init", ctypes.c_int),
                ("report_every", ctypes.c_int)]


class TrainReport(ctypes.Structure):
    _fields_ = [("loss_first", ctypes.c_float), ("loss_last", ctypes.c_float), ("loss_full", ctypes.c_float),
                ("steps", ctypes.c_int64), ("n_curve", ctypes.c_int64), ("ms", ctypes.c_double),
                ("norm", ctypes.c_float * 24)]


_LIB = None


def _sig(lib):
    H = ctypes.c_void_p
    S = [
        ("mpmx_config_default", None, [ctypes.POINTER(Config)]),
        ("mpmx_config_load_json", ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_char_p, ctypes.c_char_p]),
        ("mpmx_create", ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(H)]),
        ("mpmx_destroy", None, [H]),
        ("mpmx_last_error", ctypes.c_char_p, [H]),
        ("mpmx_status_string", ctypes.c_char_p, [ctypes.c_int]),
        ("mpmx_upload_particles", ctypes.c_int,
         [H, ctypes.c_int64, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_u8_p]),
        ("mpmx_set_objects", ctypes.c_int, [H, ctypes.POINTER(Plate), ctypes.c_int, ctypes.POINTER(Knife), ctypes.c_int]),
        ("mpmx_set_knife_keyframes", ctypes.c_int, [H, ctypes.c_int, c_float_p, ctypes.c_int, ctypes.c_float]),
        ("mpmx_set_knife_state", ctypes.c_int, [H, ctypes.c_int, ctypes.POINTER(KnifeState)]),
        ("mpmx_knife_state_at", ctypes.c_int, [H, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(KnifeState)]),
        ("mpmx_max_speed", ctypes.c_int, [H, c_float_p]),
        ("mpmx_set_dt", ctypes.c_int, [H, ctypes.c_float]),
        ("mpmx_prepare", ctypes.c_int, [H]),
        ("mpmx_get_lattice", ctypes.c_int, [H, c_int_p, c_float_p, c_float_p, c_float_p, c_float_p]),
        ("mpmx_step", ctypes.c_int, [H, ctypes.c_int]),
        ("mpmx_frame_finished", ctypes.c_int, [H, c_int_p]),
        ("mpmx_synchronize", ctypes.c_int, [H]),
        ("mpmx_download_particles", ctypes.c_int,
         [H, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_u8_p]),
        ("mpmx_download_plastic_state", ctypes.c_int, [H, c_float_p, c_float_p]),
        ("mpmx_upload_plastic_state", ctypes.c_int, [H, ctypes.c_int, c_float_p, c_float_p]),
        ("mpmx_write_particles_ply", ctypes.c_int, [H, ctypes.c_char_p]),
        ("mpmx_write_grid_ply", ctypes.c_int, [H, ctypes.c_char_p]),
        ("mpmx_checkpoint_save", ctypes.c_int, [H, ctypes.c_char_p]),
        ("mpmx_checkpoint_load", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(H)]),
        ("mpmx_pack_render", ctypes.c_int, [H, c_float_p, c_float_p, c_float_p]),
        ("mpmx_pack_render_device", ctypes.c_int, [H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        ("mpmx_get_grid", ctypes.c_int, [H, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p]),
        ("mpmx_stats_get", ctypes.c_int, [H, ctypes.POINTER(Stats)]),
        ("mpmx_profile_steps", ctypes.c_int, [H, ctypes.c_int, ctypes.POINTER(KernelTimes)]),
        ("mpmx_knife_state_eval", ctypes.c_int, [ctypes.POINTER(Knife), c_float_p, ctypes.c_int, ctypes.c_float,
                                                  ctypes.c_double, ctypes.c_int64, ctypes.POINTER(KnifeState)]),
        ("mpmx_scene_boxes", ctypes.c_int,
         [ctypes.POINTER(Box), ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_uint64, c_float_p, c_float_p,
          ctypes.c_int64, c_i64_p]),
        ("mpmx_create_batch", ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int, ctypes.POINTER(H)]),
        ("mpmx_upload_scene", ctypes.c_int,
         [H, ctypes.c_int, ctypes.c_int64, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p,
          c_float_p, c_u8_p]),
        ("mpmx_set_objects_scene", ctypes.c_int,
         [H, ctypes.c_int, ctypes.POINTER(Plate), ctypes.c_int, ctypes.POINTER(Knife), ctypes.c_int]),
        ("mpmx_scene_range", ctypes.c_int, [H, ctypes.c_int, c_i64_p, c_i64_p]),
        ("mpmx_autotune", ctypes.c_int, [H, ctypes.c_int, ctypes.POINTER(TuneReport)]),
        ("mpmx_set_launch_shape", ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        ("mpmx_closure_n_params", ctypes.c_int64, []),
        ("mpmx_closure_set_params", ctypes.c_int, [H, c_float_p, ctypes.c_int64]),
        ("mpmx_closure_get_params", ctypes.c_int, [H, c_float_p, ctypes.c_int64]),
        ("mpmx_closure_set_norm", ctypes.c_int, [H, c_float_p, c_float_p, c_float_p, c_float_p]),
        ("mpmx_harvest_begin", ctypes.c_int, [H, ctypes.c_int64]),
        ("mpmx_harvest_end", ctypes.c_int, [H, c_float_p, ctypes.c_int64, c_i64_p]),
        ("mpmx_train_config_default", None, [ctypes.POINTER(TrainConfig)]),
        ("mpmx_closure_train", ctypes.c_int, [H, c_float_p, ctypes.c_int64, ctypes.POINTER(TrainConfig),
                                              ctypes.POINTER(TrainReport), c_float_p, ctypes.c_int64]),
        ("mpmx_closure_eval", ctypes.c_int, [H, c_float_p, ctypes.c_int64, c_float_p, c_float_p, c_u8_p]),
        ("mpmx_boundary_eval", ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(Plate), ctypes.c_int,
                                              ctypes.POINTER(Knife), ctypes.c_int64, c_float_p, c_float_p,
                                              ctypes.c_int64, c_float_p]),
        ("mpmx_blob_scene", ctypes.c_int,
         [ctypes.POINTER(BlobModel), ctypes.c_int32, c_float_p, ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
          ctypes.c_uint32, ctypes.c_float, c_float_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, c_i64_p,
          c_float_p, ctypes.POINTER(BlobPlacement)]),
        ("mpmx_load_keyframes", ctypes.c_int, [ctypes.c_char_p, c_float_p, ctypes.c_int64, c_i64_p]),
        ("mpmx_mlp_forward", ctypes.c_int, [c_float_p, c_float_p, c_float_p, ctypes.c_int64, c_float_p]),
        ("mpmx_mlp_forward_noalloc", ctypes.c_int, [c_float_p, c_float_p, c_float_p, ctypes.c_int64, c_float_p]),
        ("mpmx_mlp_gradient", ctypes.c_int,
         [c_float_p, c_float_p, c_float_p, c_float_p, ctypes.c_int64, c_float_p, c_float_p]),
        ("mpmx_sample_poisson_box", ctypes.c_int, [c_float_p, c_float_p, ctypes.c_float, ctypes.c_int32,
                                                   ctypes.c_uint32, c_float_p, ctypes.c_int64, c_i64_p]),
        ("mpmx_sample_level_sets", ctypes.c_int,
         [ctypes.POINTER(LevelSet), ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.c_uint32, ctypes.c_float,
          c_float_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, c_i64_p, c_float_p]),
    ]
    for name, res, args in S:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load():
    """Load libmpmx.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pigment_mixing_amd.build`")
        _LIB = _sig(ctypes.CDLL(LIB_PATH))
    return _LIB


def fptr(a):
    return None if a is None else a.ctypes.data_as(c_float_p)


def default_config():
    c = Config()
    load().mpmx_config_default(ctypes.byref(c))
    return c


def load_config_json(sim_config_path=None, pigment_config_path=None, cfg=None):
    """Config with sim_config.json / pigment_config.json applied as the
    reference's setupScene reads them (mpmx_config_load_json)."""
    c = default_config() if cfg is None else cfg
    enc = lambda p: None if p is None else os.fsencode(p)
    rc = load().mpmx_config_load_json(ctypes.byref(c), enc(sim_config_path), enc(pigment_config_path))
    if rc:
        raise MpmxError(rc, f"mpmx_config_load_json({sim_config_path}, {pigment_config_path})")
    return c


def scene_boxes(boxes, dx, per_axis=2, seed=42):
    """Particles of a list of (lo[3], cells[3], latent[7]) boxes -> (x[n,3], pig[n,7])."""
    lib = load()
    arr = (Box * len(boxes))()
    for b, (lo, cells, lat) in zip(arr, boxes):
        b.lo[:] = [float(v) for v in lo]
        b.cells[:] = [int(v) for v in cells]
        b.latent[:] = [float(v) for v in lat]
    n = ctypes.c_int64(0)
    rc = lib.mpmx_scene_boxes(arr, len(boxes), dx, per_axis, seed, None, None, 0, ctypes.byref(n))
    if rc:
        raise RuntimeError(f"mpmx_scene_boxes: {rc}")
    x = np.zeros((n.value, 3), np.float32)
    pig = np.zeros((n.value, 7), np.float32)
    rc = lib.mpmx_scene_boxes(arr, len(boxes), dx, per_axis, seed, fptr(x), fptr(pig), n.value, ctypes.byref(n))
    if rc:
        raise RuntimeError(f"mpmx_scene_boxes: {rc}")
    return x, pig


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def train_config(**kw):
    tc = TrainConfig()
    load().mpmx_train_config_default(ctypes.byref(tc))
    for k, v in kw.items():
        setattr(tc, k, v)
    return tc


def boundary_eval(cfg, X, v, plates=(), knife=None, step=0):
    """Device boundaryCollision of grid nodes X (n, 3) with velocities v (n, 3)
    against plates and the knife at its state of `step`: the new velocities."""
    lib = load()
    Xa, va = _f32(X, (-1, 3)), _f32(v, (-1, 3))
    out = np.zeros_like(Xa)
    P = (Plate * max(1, len(plates)))(*plates) if plates else None
    K = ctypes.pointer(knife) if knife is not None else None
    rc = lib.mpmx_boundary_eval(ctypes.byref(cfg), P, len(plates), K, step, fptr(Xa), fptr(va), Xa.shape[0],
                                fptr(out))
    if rc:
        raise MpmxError(rc, "mpmx_boundary_eval")
    return out


def blob_scene(models, ratios, radius=0.01, attempts=30, seed=42, ppc=27.0):
    """The paint app's blob scene (Simulation::blobs): models are (phi[nx, ny, nz],
    voxel, origin3, volume, calibration) tuples of the blob database (ascending
    volume); returns (x (N, 3), blob index (N,), dx, placements)."""
    lib = load()
    keep, db = [], (BlobModel * len(models))()
    for i, (phi, voxel, origin, volume, calib) in enumerate(models):
        a = np.ascontiguousarray(phi, np.float32)
        keep.append(a)
        db[i].shape.phi = fptr(a)
        db[i].shape.dims[:] = list(a.shape)
        db[i].shape.voxel = voxel
        db[i].shape.origin[:] = [float(v) for v in origin]
        db[i].shape.offset[:] = [0.0, 0.0, 0.0]
        db[i].shape.scale = 1.0
        db[i].volume = volume
        db[i].calibration = calib
    r = _f32(ratios)
    place = (BlobPlacement * len(r))()
    n, dx = ctypes.c_int64(0), ctypes.c_float(0.0)
    args = (db, len(models), fptr(r), len(r), radius, attempts, seed, ppc)
    rc = lib.mpmx_blob_scene(*args, None, None, 0, ctypes.byref(n), ctypes.byref(dx), place)
    if rc not in (0, ERR_CAPACITY):
        raise MpmxError(rc, "mpmx_blob_scene")
    x = np.zeros((n.value, 3), np.float32)
    obj = np.zeros(n.value, np.int32)
    rc = lib.mpmx_blob_scene(*args, fptr(x), obj.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n.value,
                             ctypes.byref(n), ctypes.byref(dx), place)
    if rc:
        raise MpmxError(rc, "mpmx_blob_scene")
    return x, obj, float(dx.value), [place[i] for i in range(len(r))]


def load_keyframes(path):
    """Frames {pos[3], quat_xyzw[4]} of a spatula animation file (n, 7)."""
    lib = load()
    n = ctypes.c_int64(0)
    rc = lib.mpmx_load_keyframes(os.fsencode(path), None, 0, ctypes.byref(n))
    if rc not in (0, ERR_CAPACITY):
        raise MpmxError(rc, "mpmx_load_keyframes")
    fr = np.zeros((n.value, 7), np.float32)
    rc = lib.mpmx_load_keyframes(os.fsencode(path), fptr(fr), n.value, ctypes.byref(n))
    if rc:
        raise MpmxError(rc, "mpmx_load_keyframes")
    return fr


MLP_NOALLOC_MAX = 65536   # MPMX_MLP_NOALLOC_MAX


def mlp_forward(params, norm24, x8, noalloc=False):
    """Device forward pass of the closure network: y4 = net(x8) in output units.
    noalloc: through mpmx_mlp_forward_noalloc (fixed device array, at most
    MLP_NOALLOC_MAX samples)."""
    lib = load()
    p, n24, x = _f32(params), _f32(norm24), _f32(x8, (-1, 8))
    y = np.zeros((x.shape[0], 4), np.float32)
    f = lib.mpmx_mlp_forward_noalloc if noalloc else lib.mpmx_mlp_forward
    rc = f(fptr(p), fptr(n24), fptr(x), x.shape[0], fptr(y))
    if rc:
        raise MpmxError(rc, "mpmx_mlp_forward_noalloc" if noalloc else "mpmx_mlp_forward")
    return y


def mlp_gradient(params, norm24, x8, t4):
    """Device loss and parameter gradient of the closure network."""
    lib = load()
    p, n24, x, t = _f32(params), _f32(norm24), _f32(x8, (-1, 8)), _f32(t4, (-1, 4))
    g = np.zeros(p.size, np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib.mpmx_mlp_gradient(fptr(p), fptr(n24), fptr(x), fptr(t), x.shape[0], fptr(g), fptr(loss))
    if rc:
        raise MpmxError(rc, "mpmx_mlp_gradient")
    return float(loss[0]), g


def sample_poisson_box(lo, hi, radius, attempts=30, seed=42):
    """Poisson-disk points (N, 3) with minimum distance `radius` in [lo, hi]."""
    lib = load()
    lo3, hi3 = _f32(lo), _f32(hi)
    n = ctypes.c_int64(0)
    rc = lib.mpmx_sample_poisson_box(fptr(lo3), fptr(hi3), radius, attempts, seed, None, 0, ctypes.byref(n))
    if rc not in (0, ERR_CAPACITY):
        raise MpmxError(rc, "mpmx_sample_poisson_box")
    out = np.zeros((n.value, 3), np.float32)
    rc = lib.mpmx_sample_poisson_box(fptr(lo3), fptr(hi3), radius, attempts, seed, fptr(out), n.value, ctypes.byref(n))
    if rc:
        raise MpmxError(rc, "mpmx_sample_poisson_box")
    return out


def sample_level_sets(objects, radius=0.01, attempts=30, seed=42, ppc=27.0):
    """Particles of blob objects given as (phi[nx, ny, nz], voxel, origin3, offset3, scale)
    tuples: returns (x (N, 3), object index (N,), dx).  The defaults are the
    reference's call (simulation_blobs.cpp:149: radius 0.01, ppc 27); the
    scene's config must use the same ppc (particle volume dx^3 / ppc)."""
    lib = load()
    keep, sets = [], (LevelSet * len(objects))()
    for i, (phi, voxel, origin, offset, scale) in enumerate(objects):
        a = np.ascontiguousarray(phi, np.float32)
        keep.append(a)
        sets[i].phi = fptr(a)
        sets[i].dims[:] = list(a.shape)
        sets[i].voxel = voxel
        sets[i].origin[:] = [float(v) for v in origin]
        sets[i].offset[:] = [float(v) for v in offset]
        sets[i].scale = scale
    n, dx = ctypes.c_int64(0), ctypes.c_float(0.0)
    args = (sets, len(objects), radius, attempts, seed, ppc)
    rc = lib.mpmx_sample_level_sets(*args, None, None, 0, ctypes.byref(n), ctypes.byref(dx))
    if rc not in (0, ERR_CAPACITY):
        raise MpmxError(rc, "mpmx_sample_level_sets")
    x = np.zeros((n.value, 3), np.float32)
    obj = np.zeros(n.value, np.int32)
    rc = lib.mpmx_sample_level_sets(*args, fptr(x), obj.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n.value,
                                    ctypes.byref(n), ctypes.byref(dx))
    if rc:
        raise MpmxError(rc, "mpmx_sample_level_sets")
    return x, obj, float(dx.value)


class MpmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpmx error {code}: {msg}")
        self.code = code


class Handle:
    """One simulation (or a batch of scenes) on the current HIP device (thin
    wrapper over the C-ABI).  Pass a list of configs for a batched handle."""

    def __init__(self, cfg):
        self.lib = load()
        self.h = ctypes.c_void_p()
        if isinstance(cfg, (list, tuple)):
            arr = (Config * len(cfg))(*cfg)
            self._chk(self.lib.mpmx_create_batch(arr, len(cfg), ctypes.byref(self.h)), "create_batch")
            self.n_scenes = len(cfg)
            self._scene_n = [0] * len(cfg)
        else:
            self._chk(self.lib.mpmx_create(ctypes.byref(cfg), ctypes.byref(self.h)), "create")
            self.n_scenes = 1
        self.cfg = cfg
        self.n = 0

    @classmethod
    def load_checkpoint(cls, path):
        """A prepared handle continuing from a checkpoint file (mpmx_checkpoint_load)."""
        self = cls.__new__(cls)
        self.lib = load()
        self.h = ctypes.c_void_p()
        rc = self.lib.mpmx_checkpoint_load(os.fsencode(path), ctypes.byref(self.h))
        if rc != OK:
            raise MpmxError(rc, f"checkpoint_load({path})")
        self.n_scenes = 1
        self.cfg = None
        st = Stats()
        self._chk(self.lib.mpmx_stats_get(self.h, ctypes.byref(st)), "stats")
        self.n = st.n_particles
        return self

    def save_checkpoint(self, path):
        self._chk(self.lib.mpmx_checkpoint_save(self.h, os.fsencode(path)), "checkpoint_save")

    def write_particles_ply(self, path):
        self._chk(self.lib.mpmx_write_particles_ply(self.h, os.fsencode(path)), "write_particles_ply")

    def write_grid_ply(self, path):
        self._chk(self.lib.mpmx_write_grid_ply(self.h, os.fsencode(path)), "write_grid_ply")

    def _chk(self, rc, what):
        if rc != OK:
            msg = self.lib.mpmx_last_error(self.h).decode() if self.h else ""
            raise MpmxError(rc, f"{what}: {msg}")

    def close(self):
        if self.h:
            self.lib.mpmx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, x, v=None, F=None, C=None, pig=None, diff=None, eps=None, active=None):
        x = np.ascontiguousarray(x, np.float32)
        self.n = x.shape[0]
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        v, F, C, pig, diff, eps = map(conv, (v, F, C, pig, diff, eps))
        act = None if active is None else np.ascontiguousarray(active, np.uint8)
        self._chk(self.lib.mpmx_upload_particles(self.h, self.n, fptr(x), fptr(v), fptr(F), fptr(C), fptr(pig),
                                                 fptr(diff), fptr(eps),
                                                 None if act is None else act.ctypes.data_as(c_u8_p)), "upload")

    def upload_plastic_state(self, eps_vol=None, prad=None, scene=0):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        self._chk(self.lib.mpmx_upload_plastic_state(self.h, scene, fptr(conv(eps_vol)), fptr(conv(prad))),
                  "upload_plastic_state")

    def upload_scene(self, scene, x, v=None, F=None, C=None, pig=None, diff=None, eps=None, active=None):
        x = np.ascontiguousarray(x, np.float32)
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        v, F, C, pig, diff, eps = map(conv, (v, F, C, pig, diff, eps))
        act = None if active is None else np.ascontiguousarray(active, np.uint8)
        self._chk(self.lib.mpmx_upload_scene(self.h, scene, x.shape[0], fptr(x), fptr(v), fptr(F), fptr(C),
                                             fptr(pig), fptr(diff), fptr(eps),
                                             None if act is None else act.ctypes.data_as(c_u8_p)), "upload_scene")
        self._scene_n[scene] = x.shape[0]
        self.n = sum(self._scene_n)

    def set_objects_scene(self, scene, plates=(), knife=None):
        pl = (Plate * max(1, len(plates)))(*plates) if plates else None
        kn = ctypes.byref(knife) if knife is not None else None
        self._chk(self.lib.mpmx_set_objects_scene(self.h, scene, pl, len(plates), kn, 1 if knife is not None else 0),
                  "set_objects_scene")

    def scene_range(self, scene):
        a, b = ctypes.c_int64(0), ctypes.c_int64(0)
        self._chk(self.lib.mpmx_scene_range(self.h, scene, ctypes.byref(a), ctypes.byref(b)), "scene_range")
        return a.value, b.value

    def autotune(self, steps_per_trial=4):
        r = TuneReport()
        self._chk(self.lib.mpmx_autotune(self.h, steps_per_trial, ctypes.byref(r)), "autotune")
        return r

    def set_launch_shape(self, wg_p2g, wg_g2p, wg_grid):
        """Workgroups per CU for P2G, G2P and the grid update (mpmx_set_launch_shape)."""
        self._chk(self.lib.mpmx_set_launch_shape(self.h, wg_p2g, wg_g2p, wg_grid), "set_launch_shape")

    def set_objects(self, plates=(), knife=None):
        P = (Plate * max(1, len(plates)))(*plates) if plates else None
        K = ctypes.pointer(knife) if knife is not None else None
        self._chk(self.lib.mpmx_set_objects(self.h, P, len(plates), K, 1 if knife is not None else 0), "set_objects")

    def set_keyframes(self, frames7, fps=60.0, k=0):
        """Knife keyframes {pos[3], quat_xyzw[4]} per frame (MOTION_KEYFRAMES)."""
        f = _f32(frames7, (-1, 7))
        self._chk(self.lib.mpmx_set_knife_keyframes(self.h, k, fptr(f), f.shape[0], fps), "set_keyframes")

    def prepare(self):
        self._chk(self.lib.mpmx_prepare(self.h), "prepare")

    def step(self, n=1):
        self._chk(self.lib.mpmx_step(self.h, n), "step")

    def sync(self):
        self._chk(self.lib.mpmx_synchronize(self.h), "sync")

    def lattice(self):
        dims = (ctypes.c_int * 3)()
        org = np.zeros(3, np.float32)
        self._chk(self.lib.mpmx_get_lattice(self.h, dims, fptr(org), None, None, None), "lattice")
        d = list(dims)
        xs, ys, zs = (np.zeros(k, np.float32) for k in d)
        self._chk(self.lib.mpmx_get_lattice(self.h, dims, fptr(org), fptr(xs), fptr(ys), fptr(zs)), "lattice")
        return d, org, xs, ys, zs

    def download(self):
        n = self.n
        out = dict(x=np.zeros((n, 3), np.float32), v=np.zeros((n, 3), np.float32),
                   F=np.zeros((n, 3, 3), np.float32), C=np.zeros((n, 3, 3), np.float32),
                   pig=np.zeros((n, 7), np.float32), diff=np.zeros(n, np.float32),
                   eps=np.zeros(n, np.float32), dgamma=np.zeros(n, np.float32), active=np.zeros(n, np.uint8))
        self._chk(self.lib.mpmx_download_particles(
            self.h, fptr(out["x"]), fptr(out["v"]), fptr(out["F"]), fptr(out["C"]), fptr(out["pig"]),
            fptr(out["diff"]), fptr(out["eps"]), fptr(out["dgamma"]), out["active"].ctypes.data_as(c_u8_p)),
            "download")
        out["eps_vol"] = np.zeros(n, np.float32)
        out["prad"] = np.zeros(n, np.float32)
        self._chk(self.lib.mpmx_download_plastic_state(self.h, fptr(out["eps_vol"]), fptr(out["prad"])),
                  "download_plastic_state")
        return out

    def pack_render(self):
        n = self.n
        pos4 = np.zeros((n, 4), np.float32)
        lat7 = np.zeros((n, 7), np.float32)
        diff = np.zeros(n, np.float32)
        self._chk(self.lib.mpmx_pack_render(self.h, fptr(pos4), fptr(lat7), fptr(diff)), "pack_render")
        return pos4, lat7, diff

    def grid(self):
        d, _, _, _, _ = self.lattice()
        nn = d[0] * d[1] * d[2]
        out = dict(mass=np.zeros(nn, np.float32), v=np.zeros((nn, 3), np.float32),
                   flip=np.zeros((nn, 3), np.float32), pig=np.zeros((nn, 7), np.float32),
                   shear=np.zeros(nn, np.float32))
        self._chk(self.lib.mpmx_get_grid(self.h, fptr(out["mass"]), fptr(out["v"]), fptr(out["flip"]),
                                         fptr(out["pig"]), fptr(out["shear"])), "grid")
        return out

    def stats(self):
        s = Stats()
        self._chk(self.lib.mpmx_stats_get(self.h, ctypes.byref(s)), "stats")
        return s

    def profile(self, nsteps):
        kt = KernelTimes()
        self._chk(self.lib.mpmx_profile_steps(self.h, nsteps, ctypes.byref(kt)), "profile")
        return kt

    def max_speed(self):
        """Largest |v| of the live particles (mpmx_max_speed)."""
        v = ctypes.c_float(0.0)
        self._chk(self.lib.mpmx_max_speed(self.h, ctypes.byref(v)), "max_speed")
        return v.value

    def set_dt(self, dt):
        """Step of the following steps (mpmx_set_dt; objects continue from the current step)."""
        self._chk(self.lib.mpmx_set_dt(self.h, float(dt)), "set_dt")

    def knife_state(self, step):
        ks = KnifeState()
        self._chk(self.lib.mpmx_knife_state_at(self.h, 0, step, ctypes.byref(ks)), "knife_state")
        return ks

    # ---- learned closure
    def harvest_begin(self, capacity):
        self._chk(self.lib.mpmx_harvest_begin(self.h, int(capacity)), "harvest_begin")

    def harvest_end(self, max_records):
        out = np.zeros((int(max_records), 12), np.float32)
        n = ctypes.c_int64(0)
        self._chk(self.lib.mpmx_harvest_end(self.h, fptr(out), int(max_records), ctypes.byref(n)), "harvest_end")
        return out[:n.value]

    def closure_train(self, records, curve_cap=256, **kw):
        rec = _f32(records, (-1, 12))
        tc = train_config(**kw)
        rep = TrainReport()
        curve = np.zeros(curve_cap, np.float32)
        self._chk(self.lib.mpmx_closure_train(self.h, fptr(rec), rec.shape[0], ctypes.byref(tc), ctypes.byref(rep),
                                              fptr(curve), curve_cap), "closure_train")
        return rep, curve[:rep.n_curve].copy()

    def closure_params(self):
        n = self.lib.mpmx_closure_n_params()
        p = np.zeros(n, np.float32)
        self._chk(self.lib.mpmx_closure_get_params(self.h, fptr(p), n), "closure_get_params")
        return p

    def closure_eval(self, trial4):
        t = _f32(trial4, (-1, 4))
        n = t.shape[0]
        lrn = np.zeros((n, 4), np.float32)
        ex = np.zeros((n, 4), np.float32)
        acc = np.zeros(n, np.uint8)
        self._chk(self.lib.mpmx_closure_eval(self.h, fptr(t), n, fptr(lrn), fptr(ex), acc.ctypes.data_as(c_u8_p)),
                  "closure_eval")
        return lrn, ex, acc.astype(bool)

    def closure_set(self, params, norm24):
        p, n24 = _f32(params), _f32(norm24)
        self._chk(self.lib.mpmx_closure_set_norm(self.h, fptr(n24[0:8]), fptr(n24[8:16]), fptr(n24[16:20]),
                                                 fptr(n24[20:24])), "closure_set_norm")
        self._chk(self.lib.mpmx_closure_set_params(self.h, fptr(p), p.size), "closure_set_params")
