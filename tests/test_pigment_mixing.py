"""Closed-form check of the pigment mixing of G2P (g2p.cpp:98-130) together
with the pigment and shear transfers of P2G (p2g.cpp:20-120).

A block of paint in simple shear, v = A (x - x_c) with A = [[0, g, 0], 0, 0]
(APIC matrix Bmat = A dx^2 / 4, so that the transfers carry the affine field
exactly) and the shear intensity |sym L| = g / sqrt(2) as its initial
diffusion factor, holds two pigments side by side, with g = 0 and no objects.
At every step the expected latent of each particle follows from the state
before the step, computed here independently in float64:
  node latent  = sum_p w m pig / sum_p w m, first four renormalised to sum 1;
  node shear   = sum_p w m s_p / sum_p w m;
  particle     : pbar = sum_i w pig_i, sbar = sum_i w shear_i;
  mix = clamp(D_max * boost(t) * smoothstep(e0, e1, sbar) * dt, 0, 1) with
  boost(t) = 1 + clamp((t - t0) / (t1 - t0), 0, 1) (b - 1) for t > t0, else 1;
  pig' = (1 - mix) pig + mix pbar, first four renormalised to sum 1.
The edges put sbar in the cubic part of the smoothstep, the boost window
covers the checked steps and one pigment's four concentrations sum to 0.8, so
a linear smoothstep, a missing boost, a missing renormalisation (node or
particle) or an unnormalised node latent each fail the check.  The check
takes a simulator factory: here the oracle; tests/test_gpu_mixing.py runs it
on the device."""
import math

import numpy as np

from pigment_mixing_amd import abi, scenes
from tests import oracle_py as orc

DX = scenes.DX
GAMMA = 20.0               # shear rate (1/s)
SBAR = GAMMA / math.sqrt(2.0)


def mixing_scene():
    cells = (8, 6, 8)
    lo_a, lo_b = (8 * DX, 8 * DX, 8 * DX), (16 * DX, 8 * DX, 8 * DX)
    # the second latent's four pigment concentrations sum to 0.8, not 1, so
    # that the renormalisations of the node and particle latents act
    lat_b = np.array(scenes.LATENTS["hansa_yellow"], np.float64)
    lat_b[:4] *= 0.8 / lat_b[:4].sum()
    x, pig = abi.scene_boxes([(lo_a, cells, scenes.LATENTS["cobalt_blue"]), (lo_b, cells, list(lat_b))], DX, 2, 42)
    c = scenes.material_config((40, 28, 28))
    c.gravity[:] = [0.0, 0.0, 0.0]
    c.D_max = 2000.0
    c.D_edge0, c.D_edge1 = 0.5 * SBAR, 2.0 * SBAR
    c.start_boost_time, c.end_boost_time, c.boost_factor = 0.0, 8 * c.dt, 3.0
    xc = x.mean(0)
    A = np.zeros((3, 3), np.float32)
    A[0, 1] = GAMMA
    v = ((x - xc) @ A.T).astype(np.float32)
    C = np.broadcast_to(A * (DX * DX / 4.0), (x.shape[0], 3, 3)).astype(np.float32).copy()
    diff = np.full(x.shape[0], SBAR, np.float32)
    return dict(cfg=c, x=x, v=v, C=C, pig=pig, diff=diff)


def _spline(u):
    a = np.abs(u)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _stencil(x, axes):
    """Per particle: node indices (n, 27, 3) and weights (n, 27) of the quadratic
    B-spline on the lattice node coordinates `axes` (float64)."""
    n = x.shape[0]
    idx = np.zeros((n, 27, 3), np.int64)
    w = np.ones((n, 27))
    offs = np.array([(a, b, c) for a in range(3) for b in range(3) for c in range(3)])
    for d in range(3):
        ax = axes[d].astype(np.float64)
        dx = ax[1] - ax[0]
        base = np.floor((x[:, d] - ax[0]) / dx - 0.5).astype(np.int64)
        ii = base[:, None] + offs[None, :, d]
        idx[:, :, d] = ii
        w *= _spline((x[:, d][:, None] - ax[ii]) / dx)
    return idx, w


def expected_latents(state, axes, cfg, time):
    x = state["x"].astype(np.float64)
    pig = state["pig"].astype(np.float64)
    shear = state["diff"].astype(np.float64)
    idx, w = _stencil(x, axes)
    dims = [len(a) for a in axes]
    flat = (idx[..., 0] * dims[1] + idx[..., 1]) * dims[2] + idx[..., 2]
    nn = dims[0] * dims[1] * dims[2]
    m = np.bincount(flat.ravel(), w.ravel(), nn)
    gp = np.stack([np.bincount(flat.ravel(), (w * pig[:, c][:, None]).ravel(), nn) for c in range(7)], 1)
    gs = np.bincount(flat.ravel(), (w * shear[:, None]).ravel(), nn)
    has = m > 0
    gp[has] /= m[has, None]
    gs[has] /= m[has]
    s4 = gp[:, :4].sum(1)
    ok = s4 > 1e-6
    gp[ok, :4] /= s4[ok, None]
    pbar = np.einsum("pk,pkc->pc", w, gp[flat])
    sbar = np.einsum("pk,pk->p", w, gs[flat])
    t = np.clip((sbar - cfg.D_edge0) / (cfg.D_edge1 - cfg.D_edge0), 0.0, 1.0)
    ss = t * t * (3.0 - 2.0 * t)
    tf = 1.0
    if time > cfg.start_boost_time and cfg.end_boost_time > cfg.start_boost_time:
        r = min(max((time - cfg.start_boost_time) / (cfg.end_boost_time - cfg.start_boost_time), 0.0), 1.0)
        tf = 1.0 + r * (cfg.boost_factor - 1.0)
    mix = np.clip(cfg.D_max * tf * ss * cfg.dt, 0.0, 1.0)
    out = (1.0 - mix)[:, None] * pig + mix[:, None] * pbar
    s = out[:, :4].sum(1)
    ok = s > 1e-6
    out[ok, :4] /= s[ok, None]
    return out, sbar, mix


def check_pigment_mixing(make, steps=6, log=print):
    sc = mixing_scene()
    sim = make(sc["cfg"])
    sim.upload(sc["x"], v=sc["v"], C=sc["C"], pig=sc["pig"], diff=sc["diff"])
    sim.set_objects([], None)
    sim.prepare()
    _, _, xs, ys, zs = sim.lattice()
    state = sim.download()
    time = np.float32(0.0)
    moved = 0.0
    for k in range(steps):
        exp, sbar, mix = expected_latents(state, (xs, ys, zs), sc["cfg"], float(time))
        if k == 0:
            # the uploaded shear intensity is uniform: sbar is the closed form g / sqrt 2
            np.testing.assert_allclose(sbar, SBAR, rtol=1e-5)
        sim.step(1)
        new = sim.download()
        err = np.abs(new["pig"] - exp).max()
        moved = max(moved, float(np.abs(new["pig"] - state["pig"]).max()))
        log(f"[mixing] step {k}: time {float(time):.2e} mix {mix.min():.3e}..{mix.max():.3e} "
            f"max|pig - expected| {err:.2e} max change {np.abs(new['pig'] - state['pig']).max():.2e}")
        assert err <= 2e-6, (k, err)
        if k == 0:
            # the first G2P's diffusion factor is |sym L| = g / sqrt 2 where the
            # stencil sees only the affine field (1.5 dx and more inside the block)
            half = np.array([8, 3, 4]) * DX
            inner = np.all(np.abs(state["x"] - state["x"].mean(0)) < half - 1.6 * DX, axis=1)
            assert inner.sum() > 100
            np.testing.assert_allclose(new["diff"][inner], SBAR, rtol=1e-5)
        state = new
        time = np.float32(time + np.float32(sc["cfg"].dt))
    assert moved > 1e-3     # the mixing was active


def test_pigment_mixing_closed_form_on_the_oracle():
    check_pigment_mixing(lambda cfg: orc.Oracle(cfg))
