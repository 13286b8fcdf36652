"""Benchmark and parity scenes (synthetic inputs, seeded).

Each scene is a dict with: cfg (abi.Config), x, pig (particles), plates, knife,
steps, and a name.  The paint is box blobs sampled 2x2x2 per cell with seeded
uniform sub-cell jitter; latents are the reference's Mixbox latents of palette
colours (tests/golden/mixbox_latents.json holds the same numbers, written by the
reference's own Mixbox).

  paint_2blob_64x32x64   two 20x8x20-cell blobs, 51,200 particles, slab knife
                         translating at 0.1 m/s through the shared face, 2000 steps
  paint_4blob_256x64x256 four 64x8x64-cell blobs, 1,048,576 particles, slab knife
                         on a 40 mm circle at 0.15 m/s, 500 steps
  paint_6blob_512x128x512 six 102x16x102-cell blobs, 7,990,272 particles, yield
                         100 Pa, viscosity 20 Pa s, knife sweeping linearly then on a
                         40 mm circle, 300 steps
  batch_scenes           64 independent scenes (seeds 42 + s): particle counts
                         log-uniform in 20k-400k, 2-4 box blobs of random extents
                         and palette colours, cubic lattices of 64^3-160^3 nodes,
                         yield U(10, 200) Pa, viscosity U(1, 50) Pa s, slab knife
                         translating at U(0.05, 0.3) m/s, 300 steps
Material (first two): E = 1e5 Pa, nu = 0.3, rho = 1500 kg/m^3, von Mises yield 50 Pa,
viscosity 10 Pa s (perzyna_visc = 10/50 = 0.2 s), g = -9.81, flip_ratio -0.5,
pigment parameters of the reference's pigment_config.json, dt = 1e-5 s,
dx = 3.125 mm, floor SlipStick mu = 1 at y = 0, knife SlipFree mu = 0.3.
"""
import json
import math
import os

import numpy as np

from . import abi

DX = 3.125e-3

# Mixbox latents of palette colours (written by the reference's Mixbox; the same
# numbers are the fixture tests/golden/mixbox_latents.json)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "palette_latents.json")) as _f:
    LATENTS = json.load(_f)["latents"]


def material_config(dims, yield_pa=50.0, viscosity=10.0, dt=1e-5, dx=DX):
    c = abi.default_config()
    c.E = 1e5
    c.nu = 0.3
    c.rho = 1500.0
    c.flip_ratio = -0.5
    c.plastic_model = abi.PLASTIC_VMVISC
    c.q_max = c.q_min = yield_pa
    c.xi = 0.0
    c.perzyna_exp = 1.0
    c.perzyna_visc = viscosity / yield_pa
    c.gravity[:] = [0.0, -9.81, 0.0]
    # reference pigment_config.json
    c.D_max, c.D_edge0, c.D_edge1 = 50.0, 1.5, 7.0
    c.start_boost_time, c.end_boost_time, c.boost_factor = 15.0, 30.0, 3.0
    c.dx = dx
    c.ppc = 8.0
    c.grid_dims[:] = list(dims)
    c.origin_mode = 0
    c.dt = dt
    c.fps = 60.0
    c.use_musl = 0
    c.closure = abi.CLOSURE_EXACT
    c.closure_tol = 1e-4
    return c


def floor_plate(mu=1.0):
    p = abi.Plate()
    p.pos_object = 0.0
    p.plate_type = abi.PLATE_BOTTOM
    p.bc = abi.BC_SLIPSTICK
    p.friction = mu
    p.pos_lower = -1e15
    p.pos_upper = 1e15
    p.v[:] = [0.0, 0.0, 0.0]
    p.vmin_factor, p.load_factor = 1.0, 0.0
    return p


def slab_knife(motion, p0, velocity=(0, 0, 0), radius=0.0, speed=0.0, phase0=0.0, t_switch=0.0,
               half=(0.030, 0.001, 0.015), tilt_deg=30.0, mu=0.3):
    k = abi.Knife()
    k.shape = abi.KNIFE_SLAB
    k.half[:] = list(half)
    k.offset = 0.0
    k.bc = abi.BC_SLIPFREE
    k.friction = mu
    k.motion = motion
    k.p0[:] = list(p0)
    k.tilt_axis[:] = [1.0, 0.0, 0.0]
    k.tilt_angle = math.radians(tilt_deg)
    k.velocity[:] = list(velocity)
    k.radius = radius
    k.speed = speed
    k.phase0 = phase0
    k.t_switch = t_switch
    return k


def _boxes(specs, dx):
    return [((lo[0] * dx, lo[1] * dx, lo[2] * dx), cells, LATENTS[col]) for lo, cells, col in specs]


def paint_2blob(cells=(20, 8, 20), dims=(64, 32, 64), steps=2000, dx=DX, knife=True, seed=42):
    cx, cy, cz = cells
    specs = [((0, 0, 0), cells, "cobalt_blue"), ((cx, 0, 0), cells, "hansa_yellow")]
    x, pig = abi.scene_boxes(_boxes(specs, dx), dx, 2, seed)
    cfg = material_config(dims, dx=dx)
    face = cx * dx
    kn = slab_knife(abi.MOTION_LINEAR, (face - 0.015, 0.5 * cy * dx, 0.5 * cz * dx), velocity=(0.1, 0.0, 0.0)) \
        if knife else None
    return dict(name=f"paint_2blob_{dims[0]}x{dims[1]}x{dims[2]}", cfg=cfg, x=x, pig=pig,
                plates=[floor_plate()], knife=kn, steps=steps)


def paint_4blob(cells=(64, 8, 64), dims=(256, 64, 256), steps=500, dx=DX, seed=42):
    cx, cy, cz = cells
    specs = [((0, 0, 0), cells, "cobalt_blue"), ((cx, 0, 0), cells, "hansa_yellow"),
             ((0, 0, cz), cells, "cadmium_red"), ((cx, 0, cz), cells, "phthalo_green")]
    x, pig = abi.scene_boxes(_boxes(specs, dx), dx, 2, seed)
    cfg = material_config(dims, dx=dx)
    kn = slab_knife(abi.MOTION_CIRCULAR, (cx * dx, 0.5 * cy * dx, cz * dx), radius=0.040, speed=0.15)
    return dict(name=f"paint_4blob_{dims[0]}x{dims[1]}x{dims[2]}", cfg=cfg, x=x, pig=pig,
                plates=[floor_plate()], knife=kn, steps=steps)


def spatula_keyframes(p0, velocity, tilt_deg=30.0, yaw_rate=0.5, n=13, fps=60.0):
    """Synthetic knife animation in the format of the reference's
    spatula_motion.bin (object_spatula.hpp:235-250): n frames at fps of
    {pos[3], quat_xyzw[4]}.  The centre translates at `velocity`; the blade,
    tilted by tilt_deg about x, turns about y at yaw_rate rad/s."""
    frames = np.zeros((n, 7), np.float32)
    a = math.radians(tilt_deg) / 2
    qx = np.array([math.sin(a), 0.0, 0.0, math.cos(a)])
    for i in range(n):
        t = i / fps
        frames[i, :3] = np.asarray(p0) + t * np.asarray(velocity)
        b = 0.5 * yaw_rate * t
        qy = np.array([0.0, math.sin(b), 0.0, math.cos(b)])
        # q = qy * qx (Hamilton product, xyzw)
        x1, y1, z1, w1 = qy
        x2, y2, z2, w2 = qx
        frames[i, 3:] = [w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                         w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2]
    return frames


def paint_app_2blob(cells=(6, 4, 6), dims=(24, 16, 24), dx=DX, seed=42, stretch=60.0, shear=30.0):
    """The paint app's own step (simulation_units.cpp:28-186): sim_config.json's
    material (E 5e5, nu 0.2, rho 1000, friction angle 5 deg, cohesion 500 Pa,
    perzyna_visc 0.4 s) with the Drucker-Prager/Perzyna map and Pradhana's
    correction, MUSL, and the trapezoid spatula on keyframes (SlipFree).  Two
    blobs start with a velocity field of uniform expansion `stretch` (1/s) plus
    simple shear `shear` (1/s) about their centre, so that within 100 steps
    the map sees both the tension tip and the Perzyna overstress."""
    cfgdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config")
    c = abi.load_config_json(os.path.join(cfgdir, "sim_config.json"), os.path.join(cfgdir, "pigment_config.json"))
    c.plastic_model = abi.PLASTIC_DPVISC
    c.use_pradhana = 1
    c.use_musl = 1
    c.gravity[:] = [0.0, -9.81, 0.0]
    c.dx = dx
    c.ppc = 8.0
    c.grid_dims[:] = list(dims)
    c.origin_mode = 0
    c.dt = 1e-5
    c.fps = 60.0
    c.closure = abi.CLOSURE_EXACT
    cx, cy, cz = cells
    specs = [((0, 0, 0), cells, "cobalt_blue"), ((cx, 0, 0), cells, "hansa_yellow")]
    x, pig = abi.scene_boxes(_boxes(specs, dx), dx, 2, seed)
    ctr = np.array([cx * dx, 0.5 * cy * dx, 0.5 * cz * dx], np.float32)
    r = x - ctr
    v = (stretch * r).astype(np.float32)
    v[:, 0] += shear * r[:, 1]
    kn = abi.Knife()
    kn.shape = abi.KNIFE_SPATULA
    kn.half[:] = [0.006, 0.0015, 0.008]      # halfWidth, halfThickness, halfLength
    kn.offset = 0.0005
    kn.bc = abi.BC_SLIPFREE
    kn.friction = 0.0                        # simulation_units.cpp:120
    kn.motion = abi.MOTION_KEYFRAMES
    p0 = (cx * dx - 0.006, 0.5 * cy * dx, 0.5 * cz * dx)
    frames = spatula_keyframes(p0, (0.1, 0.0, 0.0))
    return dict(name=f"paint_app_2blob_{dims[0]}x{dims[1]}x{dims[2]}", cfg=c, x=x, v=v, pig=pig,
                plates=[floor_plate()], knife=kn, frames=frames, frame_fps=60.0, steps=100)


def paint_6blob(cells=(102, 16, 102), dims=(512, 128, 512), steps=300, dx=DX, seed=42):
    """The large scene: six 102x16x102-cell blobs (8 particles per cell with
    jitter, 7,990,272 particles) of six palette latents in a 3 x 2 layout on
    a 512x128x512 lattice; yield 100 Pa, viscosity 20 Pa s; the slab knife
    sweeps linearly at 0.15 m/s across the blob faces for 1.5 ms, then turns
    onto a 40 mm circle at the same speed."""
    cx, cy, cz = cells
    cols = ["cobalt_blue", "hansa_yellow", "cadmium_red", "phthalo_green", "ultramarine_blue", "cadmium_orange"]
    specs = [((ix * cx, 0, iz * cz), cells, cols[3 * iz + ix]) for iz in range(2) for ix in range(3)]
    x, pig = abi.scene_boxes(_boxes(specs, dx), dx, 2, seed)
    cfg = material_config(dims, yield_pa=100.0, viscosity=20.0, dx=dx)
    kn = slab_knife(abi.MOTION_LINEAR_THEN_CIRCULAR, (cx * dx - 0.02, 0.5 * cy * dx, cz * dx), velocity=(0.15, 0.0, 0.0),
                    radius=0.040, speed=0.15, phase0=0.0, t_switch=1.5e-3)
    return dict(name=f"paint_6blob_{dims[0]}x{dims[1]}x{dims[2]}", cfg=cfg, x=x, pig=pig,
                plates=[floor_plate()], knife=kn, steps=steps)


def batch_scene(s, seed=42, np_range=(20_000, 400_000), dx=DX):
    """Scene s of the batched workload (its own seed 42 + s)."""
    rng = np.random.default_rng(seed + s)
    n_target = math.exp(rng.uniform(math.log(np_range[0]), math.log(np_range[1])))
    k = int(rng.integers(2, 5))                       # 2..4 blobs
    share = rng.dirichlet(np.full(k, 4.0))            # uneven blob sizes
    cols = math.ceil(math.sqrt(k))
    cy = 8
    while True:
        sides = [max(4, int(round(math.sqrt(n_target * f / 8.0 / cy)))) for f in share]
        extent = cols * max(sides)
        if extent + 16 <= 160 or cy >= 32:
            break
        cy += 4
    g = int(min(160, max(64, extent + 16)))
    names = sorted(LATENTS)
    specs, offs = [], 4
    for b, side in enumerate(sides):
        ci, cj = b % cols, b // cols
        lo = (offs + ci * max(sides), 0, offs + cj * max(sides))
        specs.append((lo, (side, cy, side), names[int(rng.integers(len(names)))]))
    x, pig = abi.scene_boxes(_boxes(specs, dx), dx, 2, seed + s)
    yield_pa = float(rng.uniform(10.0, 200.0))
    visc = float(rng.uniform(1.0, 50.0))
    speed = float(rng.uniform(0.05, 0.3))
    cfg = material_config((g, g, g), yield_pa=yield_pa, viscosity=visc, dx=dx)
    x_mid = (offs + 0.5 * max(sides)) * dx
    kn = slab_knife(abi.MOTION_LINEAR, (x_mid - 0.015, 0.5 * cy * dx, (offs + 0.5 * max(sides)) * dx),
                    velocity=(speed, 0.0, 0.0))
    return dict(name=f"batch_{s}_{g}^3", cfg=cfg, x=x, pig=pig, plates=[floor_plate()], knife=kn, steps=300,
                yield_pa=yield_pa, viscosity=visc, knife_speed=speed)


def batch_scenes(n=64, seed=42):
    return [batch_scene(s, seed) for s in range(n)]


def make_batch_handle(scene_list):
    h = abi.Handle([sc["cfg"] for sc in scene_list])
    for i, sc in enumerate(scene_list):
        h.upload_scene(i, sc["x"], v=sc.get("v"), pig=sc["pig"])
        h.set_objects_scene(i, sc["plates"], sc["knife"])
        if sc.get("frames") is not None:
            h.set_keyframes(sc["frames"], sc.get("frame_fps", 60.0), k=i)
    h.prepare()
    return h


def make_handle(scene):
    h = abi.Handle(scene["cfg"])
    h.upload(scene["x"], v=scene.get("v"), pig=scene["pig"])
    h.set_objects(scene["plates"], scene["knife"])
    if scene.get("frames") is not None:
        h.set_keyframes(scene["frames"], scene.get("frame_fps", 60.0))
    h.prepare()
    return h
