"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so) for the tests.

The oracle is the checker only: tests compare the HIP path (libmpmx.so, through
the C-ABI) against it on the same seeded inputs.
"""
import ctypes
import os
import subprocess

import numpy as np

from pigment_mixing_amd import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "oracle", "_build", "liboracle.so")
LIB_FAST = os.path.join(REPO, "oracle", "_build", "liboracle_fast.so")

F = abi.c_float_p
H = ctypes.c_void_p


def _load(path):
    # `make` is a no-op when the build is current; on a box without make or
    # sources the prebuilt library is used as is.
    if os.path.exists(os.path.join(REPO, "oracle", "Makefile")):
        try:
            subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all"], check=True,
                           capture_output=True, timeout=300)
        except (OSError, subprocess.SubprocessError):
            if not os.path.exists(path):
                raise
    lib = ctypes.CDLL(path)
    sig = [
        ("orc_create", H, [ctypes.POINTER(abi.Config)]),
        ("orc_destroy", None, [H]),
        ("orc_upload", ctypes.c_int, [H, ctypes.c_int64, F, F, F, F, F, F, F, abi.c_u8_p]),
        ("orc_set_objects", ctypes.c_int, [H, ctypes.POINTER(abi.Plate), ctypes.c_int, ctypes.POINTER(abi.Knife),
                                           ctypes.c_int]),
        ("orc_set_keyframes", ctypes.c_int, [H, F, ctypes.c_int, ctypes.c_float]),
        ("orc_prepare", ctypes.c_int, [H]),
        ("orc_step", ctypes.c_int, [H, ctypes.c_int]),
        ("orc_set_dt", ctypes.c_int, [H, ctypes.c_float]),
        ("orc_download", ctypes.c_int, [H, F, F, F, F, F, F, F, F, abi.c_u8_p]),
        ("orc_get_lattice", ctypes.c_int, [H, abi.c_int_p, F, F, F, F]),
        ("orc_get_grid", ctypes.c_int, [H, F, F, F, F, F]),
        ("orc_stats", ctypes.c_int, [H, abi.c_i64_p, abi.c_int_p, abi.c_i64_p, ctypes.POINTER(ctypes.c_double)]),
        ("orc_knife_state_at", ctypes.c_int, [ctypes.POINTER(abi.Knife), F, ctypes.c_int, ctypes.c_float,
                                              ctypes.c_double, ctypes.c_int64, F, F, F, F, F]),
        ("orc_principal", None, [F, F, F, F, F]),
        ("orc_boundary_collision", ctypes.c_int, [ctypes.POINTER(abi.Config), ctypes.POINTER(abi.Plate), ctypes.c_int,
                                                  ctypes.POINTER(abi.Knife), ctypes.c_int64, F, F, F]),
        ("orc_spline_N", ctypes.c_float, [ctypes.c_float]),
        ("orc_spline_dN", ctypes.c_float, [ctypes.c_float]),
        ("orc_knife_sdf", ctypes.c_float, [ctypes.c_int, F, ctypes.c_float, F]),
        ("orc_return_map", ctypes.c_int, [ctypes.POINTER(abi.Config), ctypes.c_int64, F, F, F, F, F,
                                          ctypes.POINTER(ctypes.c_int32)]),
        ("orc_kirchhoff", None, [ctypes.POINTER(abi.Config), F, F]),
        ("orc_return_map_state", ctypes.c_int, [ctypes.POINTER(abi.Config), ctypes.c_int64, F, F, F, F,
                                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        ("orc_download_vol", ctypes.c_int, [H, F, F]),
        ("orc_upload_vol", ctypes.c_int, [H, F, F]),
        ("orc_omp_max_threads", ctypes.c_int, []),
    ]
    for name, res, args in sig:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIBS = {}


def lib(fast=False):
    p = LIB_FAST if fast else LIB
    if p not in _LIBS:
        _LIBS[p] = _load(p)
    return _LIBS[p]


def fp(a):
    return abi.fptr(a)


class Oracle:
    def __init__(self, cfg, fast=False):
        self.lib = lib(fast)
        self.cfg = cfg
        self.h = self.lib.orc_create(ctypes.byref(cfg))
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def upload(self, x, v=None, F_=None, C=None, pig=None, diff=None, eps=None, active=None):
        x = np.ascontiguousarray(x, np.float32)
        self.n = x.shape[0]
        c = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        v, F_, C, pig, diff, eps = map(c, (v, F_, C, pig, diff, eps))
        act = None if active is None else np.ascontiguousarray(active, np.uint8)
        self._keep = (x, v, F_, C, pig, diff, eps, act)
        self.lib.orc_upload(self.h, self.n, fp(x), fp(v), fp(F_), fp(C), fp(pig), fp(diff), fp(eps),
                            None if act is None else act.ctypes.data_as(abi.c_u8_p))

    def upload_vol(self, eps_vol=None, prad=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        self.lib.orc_upload_vol(self.h, fp(conv(eps_vol)), fp(conv(prad)))

    def set_objects(self, plates=(), knife=None):
        P = (abi.Plate * max(1, len(plates)))(*plates) if plates else None
        K = ctypes.pointer(knife) if knife is not None else None
        self.lib.orc_set_objects(self.h, P, len(plates), K, 1 if knife is not None else 0)

    def set_keyframes(self, frames7, fps=60.0):
        f = np.ascontiguousarray(frames7, np.float32)
        self.lib.orc_set_keyframes(self.h, fp(f), f.shape[0], fps)

    def prepare(self):
        self.lib.orc_prepare(self.h)

    def step(self, n=1):
        return self.lib.orc_step(self.h, n)

    def set_dt(self, dt):
        """Step of the following steps (scenes without moving objects)."""
        assert self.lib.orc_set_dt(self.h, float(dt)) == 0

    def download(self):
        n = self.n
        out = dict(x=np.zeros((n, 3), np.float32), v=np.zeros((n, 3), np.float32),
                   F=np.zeros((n, 3, 3), np.float32), C=np.zeros((n, 3, 3), np.float32),
                   pig=np.zeros((n, 7), np.float32), diff=np.zeros(n, np.float32),
                   eps=np.zeros(n, np.float32), dgamma=np.zeros(n, np.float32), active=np.zeros(n, np.uint8))
        self.lib.orc_download(self.h, fp(out["x"]), fp(out["v"]), fp(out["F"]), fp(out["C"]), fp(out["pig"]),
                              fp(out["diff"]), fp(out["eps"]), fp(out["dgamma"]),
                              out["active"].ctypes.data_as(abi.c_u8_p))
        out["eps_vol"] = np.zeros(n, np.float32)
        out["prad"] = np.zeros(n, np.float32)
        self.lib.orc_download_vol(self.h, fp(out["eps_vol"]), fp(out["prad"]))
        return out

    def lattice(self):
        dims = (ctypes.c_int * 3)()
        org = np.zeros(3, np.float32)
        self.lib.orc_get_lattice(self.h, dims, fp(org), None, None, None)
        d = list(dims)
        xs, ys, zs = (np.zeros(k, np.float32) for k in d)
        self.lib.orc_get_lattice(self.h, dims, fp(org), fp(xs), fp(ys), fp(zs))
        return d, org, xs, ys, zs

    def grid(self):
        d = self.lattice()[0]
        nn = d[0] * d[1] * d[2]
        out = dict(mass=np.zeros(nn, np.float32), v=np.zeros((nn, 3), np.float32), flip=np.zeros((nn, 3), np.float32),
                   pig=np.zeros((nn, 7), np.float32), shear=np.zeros(nn, np.float32))
        self.lib.orc_get_grid(self.h, fp(out["mass"]), fp(out["v"]), fp(out["flip"]), fp(out["pig"]),
                              fp(out["shear"]))
        return out

    def stats(self):
        pc = ctypes.c_int64()
        ex = ctypes.c_int()
        na = ctypes.c_int64()
        t = ctypes.c_double()
        self.lib.orc_stats(self.h, ctypes.byref(pc), ctypes.byref(ex), ctypes.byref(na), ctypes.byref(t))
        return dict(plastic_count=pc.value, exit=ex.value, n_active=na.value, time=t.value)


def from_scene(scene, fast=False, n_threads=1):
    cfg = abi.Config.from_buffer_copy(scene["cfg"])
    cfg.n_threads = n_threads
    o = Oracle(cfg, fast=fast)
    o.upload(scene["x"], v=scene.get("v"), pig=scene["pig"])
    o.set_objects(scene["plates"], scene["knife"])
    if scene.get("frames") is not None:
        o.set_keyframes(scene["frames"], scene.get("frame_fps", 60.0))
    o.prepare()
    return o


def principal(A):
    """principal_of(F): U, sigma, V (F = U diag(sigma) V^T) and lam = sigma^2 - 1."""
    A = np.ascontiguousarray(A, np.float32).reshape(9)
    U = np.zeros(9, np.float32)
    s = np.zeros(3, np.float32)
    V = np.zeros(9, np.float32)
    lam = np.zeros(3, np.float32)
    lib().orc_principal(fp(A), fp(U), fp(s), fp(V), fp(lam))
    return U.reshape(3, 3), s, V.reshape(3, 3), lam


def boundary_collision(cfg, X, v, plates=(), knife=None, step=0):
    """Velocity of a grid node at X after boundaryCollision (one node)."""
    P = (abi.Plate * max(1, len(plates)))(*plates) if plates else None
    K = ctypes.pointer(knife) if knife is not None else None
    X = np.ascontiguousarray(X, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.zeros(3, np.float32)
    rc = lib().orc_boundary_collision(ctypes.byref(cfg), P, len(plates), K, step, fp(X), fp(v), fp(out))
    assert rc == 0
    return out


def return_map(cfg, Ft, eps=None):
    Ft = np.ascontiguousarray(Ft, np.float32).reshape(-1, 3, 3)
    n = Ft.shape[0]
    e = np.zeros(n, np.float32) if eps is None else np.ascontiguousarray(eps, np.float32)
    Fo = np.zeros_like(Ft)
    eo = np.zeros(n, np.float32)
    dg = np.zeros(n, np.float32)
    pl = np.zeros(n, np.int32)
    rc = lib().orc_return_map(ctypes.byref(cfg), n, fp(Ft), fp(e), fp(Fo), fp(eo), fp(dg),
                              pl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return Fo, eo, dg, pl, rc


def return_map_state(cfg, Ft, state3=None):
    """The configured return map with the whole plastic state: state3 rows are
    (eps_pl_dev, eps_pl_vol, eps_pl_vol_pradhana).  Returns F, out4 rows
    (the three after the map, delta_gamma), changed (F replaced), counted
    (plastic_count increment) and the status."""
    Ft = np.ascontiguousarray(Ft, np.float32).reshape(-1, 3, 3)
    n = Ft.shape[0]
    st = np.zeros((n, 3), np.float32) if state3 is None else np.ascontiguousarray(state3, np.float32).reshape(n, 3)
    Fo = np.zeros_like(Ft)
    out4 = np.zeros((n, 4), np.float32)
    ch = np.zeros(n, np.int32)
    cnt = np.zeros(n, np.int32)
    i32p = ctypes.POINTER(ctypes.c_int32)
    rc = lib().orc_return_map_state(ctypes.byref(cfg), n, fp(Ft), fp(st), fp(Fo), fp(out4),
                                    ch.ctypes.data_as(i32p), cnt.ctypes.data_as(i32p))
    return Fo, out4, ch, cnt, rc


def kirchhoff(cfg, F_):
    F_ = np.ascontiguousarray(F_, np.float32).reshape(9)
    t = np.zeros(9, np.float32)
    lib().orc_kirchhoff(ctypes.byref(cfg), fp(F_), fp(t))
    return t.reshape(3, 3)


def knife_state(knife, dt, step, frames=None, fps=60.0):
    T, Ti, N, v, w = (np.zeros(k, np.float32) for k in (12, 12, 9, 3, 3))
    fr = None if frames is None else np.ascontiguousarray(frames, np.float32)
    lib().orc_knife_state_at(ctypes.byref(knife), fp(fr), 0 if fr is None else fr.shape[0], fps, dt, step,
                             fp(T), fp(Ti), fp(N), fp(v), fp(w))
    return dict(T=T, Tinv=Ti, N=N, v=v, w=w)
