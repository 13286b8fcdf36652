"""Scenes and analytic checks of the reference's own physics tests
(matter/tests/tests.cpp), shared by the oracle tests (CPU) and the GPU tests.

The reference runs them with adaptive dt and a re-meshing lattice for seconds
to minutes of simulated time; here each runs on a fixed lattice with a fixed
dt for a shorter time, with the analytic quantity of the original test:
  incline  CoulombFrictionTest.{Plate,General}{SlipFree,SlipStick} (:100-306):
           mean displacement 0.5 g t^2 (sin th - mu cos th) of an elastic block
           on a plate or on a general object (the slab knife as the ground)
           under gravity tilted 24 deg, mu = tan 15 deg;
  slide    BoundaryTest.Analytic{SlipFree,SlipStick} (:13-98): a particle
           sliding without friction down a curve reaches sqrt(2 g h); here down
           a slab knife tilted 30 deg;
  bulk     ElasticityTest.BulkModulus (:475-540): a block squeezed between two
           NoSlip plates moving towards each other (slowly during the first
           frame, vmin_factor 10): volume-averaged Kirchhoff pressure over
           -log(J_avg) (save_data.cpp:259-288) is the bulk modulus within 3 %;
  rotation EnergyTest.Rotation (:542-602): a block spinning and translating
           with g = 0 loses kinetic energy, by at most 1e-3 of it, and never
           gains any.
A scene is a dict (cfg, x, v, C, pig, plates, knife, steps); `run` steps it on
a simulator made by `make(cfg)` (the oracle or a device handle)."""
import math

import numpy as np

from pigment_mixing_amd import abi, scenes

G = 9.81


def run(sc, make, steps=None):
    sim = make(sc["cfg"])
    sim.upload(sc["x"], v=sc.get("v"), C=sc.get("C"), pig=sc["pig"])
    sim.set_objects(sc["plates"], sc["knife"])
    sim.prepare()
    sim.step(sc["steps"] if steps is None else steps)
    return sim, sim.download()


def _elastic_cfg(dims, dx, E=1e6, rho=1000.0, dt=2e-4, flip=-0.95):
    c = scenes.material_config(dims, dx=dx, dt=dt)
    c.plastic_model = abi.PLASTIC_NONE
    c.E, c.nu, c.rho = E, 0.3, rho
    c.flip_ratio = flip
    return c


def _block(lo, cells, dx, seed=42):
    return abi.scene_boxes([(lo, cells, scenes.LATENTS["cobalt_blue"])], dx, 2, seed)


# ------------------------------------------------------------------ incline
THETA = math.radians(24.0)
MU = math.tan(math.radians(15.0))


def incline_scene(ground="plate", bc=abi.BC_SLIPFREE, t_end=0.2):
    """The block of the reference test slides along +x for t_end seconds; the
    lattice is long enough for it (dx 1 cm)."""
    dx = 0.01
    dt = 2e-4
    steps = int(round(t_end / dt))
    slide = 0.5 * G * t_end ** 2 * (math.sin(THETA) - MU * math.cos(THETA))
    nx = max(48, int(math.ceil((slide + 0.2) / dx)) + 8)
    y0 = 0.0 if ground == "plate" else 0.05
    x, pig = _block((0.0, y0 + 0.5 * dx, 0.0), (8, 3, 4), dx)
    c = _elastic_cfg((nx, 12, 12), dx, dt=dt)
    c.gravity[:] = [G * math.sin(THETA), -G * math.cos(THETA), 0.0]
    plates, knife = [], None
    if ground == "plate":
        p = scenes.floor_plate(MU)
        p.bc = bc
        plates = [p]
    else:
        # the ground is the slab knife (a general object, boundary_collision.cpp:30-118):
        # top face at y0, static, far above the thin-film layer y < 0.8 dx
        hx = 0.5 * nx * dx + 0.1
        knife = scenes.slab_knife(abi.MOTION_STATIC, (hx - 0.1, y0 - 0.02, 0.02), half=(hx, 0.02, 0.1), tilt_deg=0.0,
                                  mu=MU)
        knife.bc = bc
    return dict(cfg=c, x=x, pig=pig, plates=plates, knife=knife, steps=steps)


def incline_error(sc, out):
    t = sc["steps"] * sc["cfg"].dt
    ref = 0.5 * G * t * t * (math.sin(THETA) - MU * math.cos(THETA))
    got = float((out["x"][:, 0].astype(np.float64) - sc["x"][:, 0]).mean())
    return got, ref, abs(got - ref) / ref


# ------------------------------------------------------------------ slide
SLOPE = math.radians(30.0)


def slide_scene(bc=abi.BC_SLIPFREE, t_end=0.2):
    """One particle released just above a frictionless slab tilted 30 deg about
    z (normal (sin 30, cos 30, 0): it slides towards +x and down) for t_end
    seconds (dx 2.5 mm, dt 2e-5: a lone particle needs dt below dx / c)."""
    dx = 0.0025
    dt = 2e-5
    steps = int(round(t_end / dt))
    run_len = 0.5 * G * math.sin(SLOPE) * t_end ** 2 + 0.05
    tang = np.array([math.cos(SLOPE), -math.sin(SLOPE), 0.0])
    n = np.array([math.sin(SLOPE), math.cos(SLOPE), 0.0])
    half = (0.5 * run_len + 0.03, 0.01, 0.01)
    nx = int(math.ceil((2 * half[0] * math.cos(SLOPE) + 0.04) / dx)) + 8
    ny = int(math.ceil((2 * half[0] * math.sin(SLOPE) + 0.08) / dx)) + 8
    ctr = np.array([0.5 * nx * dx, 0.5 * ny * dx, 0.01])
    c = _elastic_cfg((nx, ny, 8), dx, dt=dt, flip=-0.95)
    c.ppc = 1.0                                   # particle volume dx^3, as the reference test
    c.gravity[:] = [0.0, -G, 0.0]
    c.origin_mode = 1
    c.grid_origin[:] = [0.0, 0.0, 0.0]
    knife = scenes.slab_knife(abi.MOTION_STATIC, tuple(ctr), half=half, tilt_deg=-30.0, mu=0.0)
    knife.tilt_axis[:] = [0.0, 0.0, 1.0]
    knife.bc = bc
    x0 = ctr + n * (half[1] + 0.5 * dx) - tang * (half[0] - 0.02)
    x = x0[None].astype(np.float32)
    pig = np.array([scenes.LATENTS["cobalt_blue"]], np.float32)
    return dict(cfg=c, x=x, pig=pig, plates=[], knife=knife, steps=steps)


def slide_error(sc, out):
    drop = float(sc["x"][0, 1] - out["x"][0, 1])
    v = float(np.linalg.norm(out["v"][0].astype(np.float64)))
    ref = math.sqrt(2 * G * drop)
    return v, ref, abs(v - ref) / ref, drop


# ------------------------------------------------------------------ bulk modulus
def bulk_scene(cells=(20, 20, 8), steps=1000):
    dx = 0.0025
    x, pig = _block((0.0, 0.0, 0.0), cells, dx)
    c = _elastic_cfg((cells[0] + 12, cells[1] + 12, cells[2] + 8), dx, rho=10000.0, dt=1e-4)
    c.gravity[:] = [0.0, 0.0, 0.0]
    c.fps = 100.0                                  # first frame (10 ms) at vel / vmin_factor
    vel = 2.5e-3 * cells[1] / 20.0                 # 1 % of the height per 0.1 s
    ly = cells[1] * dx
    bottom = scenes.floor_plate(0.0)
    bottom.bc = abi.BC_NOSLIP
    bottom.pos_object = -0.5 * dx
    bottom.v[:] = [0.0, vel, 0.0]
    top = scenes.floor_plate(0.0)
    top.bc = abi.BC_NOSLIP
    top.plate_type = abi.PLATE_TOP
    top.pos_object = ly + 0.5 * dx
    top.v[:] = [0.0, -vel, 0.0]
    for p in (bottom, top):
        p.vmin_factor, p.load_factor = 10.0, 1.0
    return dict(cfg=c, x=x, pig=pig, plates=[bottom, top], knife=None, steps=steps)


def hencky_tau(cfg, F):
    """Kirchhoff stress of F (n, 3, 3), float64: 2 mu ln V + lambda tr(ln V) I."""
    E, nu = cfg.E, cfg.nu
    mu = E / (2 * (1 + nu))
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    F = F.astype(np.float64)
    w, Q = np.linalg.eigh(F @ np.swapaxes(F, 1, 2))
    lnV = np.einsum("nij,nj,nkj->nik", Q, 0.5 * np.log(w), Q)
    tr = np.trace(lnV, axis1=1, axis2=2)
    return 2 * mu * lnV + lam * tr[:, None, None] * np.eye(3)


def bulk_modulus(cfg, F, p=None):
    """computeAvgData (save_data.cpp:259-288) and the test's K = p / -log(J_avg);
    p per particle from the caller (e.g. the device's stored stress) or from F."""
    J = np.linalg.det(F.astype(np.float64))
    if p is None:
        p = -np.trace(hencky_tau(cfg, F), axis1=1, axis2=2) / 3.0
    p_avg = float((p * J).sum() / J.sum())
    J_avg = float(J.mean())
    K_true = cfg.nu * cfg.E / ((1 + cfg.nu) * (1 - 2 * cfg.nu)) + 2.0 * cfg.E / (2 * (1 + cfg.nu)) / 3.0
    K = p_avg / -math.log(J_avg)
    return K, K_true, abs(K - K_true) / K_true, J_avg


# ------------------------------------------------------------------ rotation
def rotation_scene(cells=(20, 20, 5), steps=2000, omega=5.0):
    dx = 0.01
    x, pig = _block((0.0, 0.0, 0.0), cells, dx)
    t_end = steps * 1e-4
    margin = int(math.ceil((0.5 * t_end + 0.3 * cells[0] * dx) / dx)) + 4
    c = _elastic_cfg((cells[0] + 2 * margin, cells[1] + 2 * margin, cells[2] + 8), dx, rho=1550.0, dt=1e-4, flip=-1.0)
    c.gravity[:] = [0.0, 0.0, 0.0]
    c.origin_mode = 1
    c.grid_origin[:] = [-margin * dx, -margin * dx, -4 * dx]
    xc = x.mean(0)
    v = np.zeros_like(x)
    v[:, 0] = -omega * (x[:, 1] - xc[1]) + 0.5
    v[:, 1] = omega * (x[:, 0] - xc[0]) + 0.5
    return dict(cfg=c, x=x, v=v.astype(np.float32), pig=pig, plates=[], knife=None, steps=steps)


def kinetic(v):
    return 0.5 * float((v.astype(np.float64) ** 2).sum())
