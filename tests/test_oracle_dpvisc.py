"""The oracle's Drucker-Prager/Perzyna return map with Pradhana's correction
(plasticity.cpp:240-340), pinned by what the mechanics fixes rather than by
re-evaluating the map's own formulas:

  * tension beyond the cohesion tip lands exactly on the tip: isotropic strain,
    p = p_tip = -c/M, and the volume gain (p_tip - p_trial)/K enters eps_pl_vol
    and the Pradhana accumulator;
  * in the rate-independent limit the returned state lies on the yield surface
    q = M p + c, keeps its pressure (the flow is deviatoric) and its deviatoric
    direction;
  * with viscosity the returned state satisfies the Perzyna consistency
    condition q (dt / (visc dgam + dt))^n = q_y (closed form n = 1, Newton for
    n = 2), with dgam read back from eps_pl_dev;
  * the Pradhana shift -K eps_pl_vol_pradhana moves the tip decision, and the
    accumulator resets on every state right of the shifted tip;
  * isotropy: F -> R F gives F' -> R F'.
Every check computes strains with a float64 SVD of the oracle's FP32 input and
output, independently of the oracle's own Jacobi SVD.
"""
import math

import numpy as np

from pigment_mixing_amd import abi, scenes
from tests import oracle_py as orc

QP = 1 / math.sqrt(2)


def dp_cfg(visc=0.0, cohesion=500.0, angle_deg=5.0, exp=1.0, pradhana=1, E=5e5, nu=0.2, dt=1e-5):
    c = scenes.material_config((8, 8, 8))
    c.E, c.nu = E, nu
    c.plastic_model = abi.PLASTIC_DPVISC
    c.M = math.tan(math.radians(angle_deg))
    c.q_cohesion = cohesion
    c.perzyna_visc = visc
    c.perzyna_exp = exp
    c.use_pradhana = pradhana
    c.dt = dt
    return c


def moduli(c):
    mu = c.E / (2 * (1 + c.nu))
    lam = c.nu * c.E / ((1 + c.nu) * (1 - 2 * c.nu))
    K = lam + 2 * mu / 3
    return mu, K, 2 * QP * mu, 2 * QP * QP * mu


def hencky(F):
    return np.log(np.linalg.svd(np.asarray(F, np.float64), compute_uv=False))


def p_q(c, F):
    _, K, e_mu, _ = moduli(c)
    h = hencky(F)
    dev = h - h.mean()
    return -K * h.sum(), e_mu * np.linalg.norm(dev), dev


def rng(s):
    return np.random.default_rng(s)


def rand_rot(r):
    Q, _ = np.linalg.qr(r.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def trial_F(r, vol, dev_amp):
    """F = R1 diag(e^h) R2 with tr h = vol and a random deviator of size dev_amp."""
    d = r.standard_normal(3)
    d -= d.mean()
    d *= dev_amp / np.linalg.norm(d)
    h = d + vol / 3
    return (rand_rot(r) @ np.diag(np.exp(h)) @ rand_rot(r)).astype(np.float32)


def test_tension_beyond_tip_projects_to_tip():
    c = dp_cfg()
    _, K, _, _ = moduli(c)
    p_tip = -c.q_cohesion / c.M
    r = rng(11)
    Ft = np.stack([trial_F(r, vol=r.uniform(0.025, 0.05), dev_amp=r.uniform(0, 3e-3)) for _ in range(64)])
    st = np.zeros((64, 3), np.float32)
    st[:, 0] = r.uniform(0, 1e-2, 64)       # eps_pl_dev
    st[:, 1] = r.uniform(-1e-3, 1e-3, 64)   # eps_pl_vol
    Fo, out, ch, cnt, rc = orc.return_map_state(c, Ft, st)
    assert rc == 0 and ch.all() and cnt.all()
    for k in range(64):
        p0, q0, _ = p_q(c, Ft[k])
        assert p0 < p_tip
        h1 = hencky(Fo[k])
        # isotropic (q = 0) to the FP32 resolution of the rotated product U D V^T
        np.testing.assert_allclose(h1, h1.mean(), atol=5e-7)
        # on the tip, to the FP32 resolution of F = U D V^T (U, V orthonormal to
        # ~2e-7 per column: up to ~1.2e-6 in tr h, times K)
        np.testing.assert_allclose(-K * h1.sum(), p_tip, rtol=0, atol=K * 2e-6)
        inst = (p_tip - p0) / K                                         # volume gained
        np.testing.assert_allclose(out[k, 1], st[k, 1] + inst, atol=2e-6)
        np.testing.assert_allclose(out[k, 2], st[k, 2] + inst, atol=2e-6)
        # the whole deviator is removed: eps_pl_dev grows by |dev h| = q / e_mu
        np.testing.assert_allclose(out[k, 0], st[k, 0] + q0 / moduli(c)[2], atol=2e-6)
        np.testing.assert_allclose(out[k, 3], q0 / moduli(c)[2] / QP / c.dt, rtol=1e-3, atol=0.3)


def test_rate_independent_limit_lands_on_yield_surface():
    c = dp_cfg(visc=0.0)
    r = rng(12)
    Ft = np.stack([trial_F(r, vol=r.uniform(-2e-2, 1e-2), dev_amp=r.uniform(2e-3, 1e-2)) for _ in range(200)])
    Fo, out, ch, cnt, rc = orc.return_map_state(c, Ft)
    assert rc == 0
    n_pl = 0
    for k in range(200):
        p0, q0, d0 = p_q(c, Ft[k])
        q_y = c.M * p0 + c.q_cohesion
        if q0 <= q_y * (1 - 1e-4):
            assert not ch[k]
            np.testing.assert_array_equal(Fo[k], Ft[k])
            continue
        if q0 < q_y * (1 + 1e-4):
            continue                                              # FP32 decision band
        n_pl += 1
        assert ch[k] and cnt[k] == 1
        p1, q1, d1 = p_q(c, Fo[k])
        np.testing.assert_allclose(p1, p0, rtol=1e-4, atol=1.0)   # deviatoric flow keeps the pressure
        np.testing.assert_allclose(q1, q_y, rtol=2e-3)            # on q = M p + c
        assert np.dot(d0, d1) / (np.linalg.norm(d0) * np.linalg.norm(d1)) > 1 - 1e-4
        assert out[k, 2] == 0.0                                   # right of the tip: accumulator reset
    assert n_pl > 100


def _consistency(c, Ft, Fo, out, st):
    """Perzyna consistency q_new (dt/(visc dgam + dt))^n - q_y for every
    projected state, relative to q_y; dgam from the eps_pl_dev increment."""
    res = []
    for k in range(len(Ft)):
        p0, q0, _ = p_q(c, Ft[k])
        q_y = c.M * p0 + c.q_cohesion
        if q0 < q_y * (1 + 1e-3):
            continue
        _, q1, _ = p_q(c, Fo[k])
        dgam = (float(out[k, 0]) - float(st[k, 0])) / QP
        assert dgam > 0
        res.append((q1 * (c.dt / (c.perzyna_visc * dgam + c.dt)) ** c.perzyna_exp - q_y) / q_y)
        np.testing.assert_allclose(out[k, 3], dgam / c.dt, rtol=1e-3)
    return np.array(res)


def test_perzyna_consistency_closed_form_and_newton():
    r = rng(13)
    Ft = np.stack([trial_F(r, vol=r.uniform(-1e-2, 5e-3), dev_amp=r.uniform(3e-3, 2e-2)) for _ in range(200)])
    st = np.zeros((200, 3), np.float32)
    for exp in (1.0, 2.0):
        c = dp_cfg(visc=0.4, exp=exp)
        Fo, out, ch, cnt, rc = orc.return_map_state(c, Ft, st)
        assert rc == 0
        res = _consistency(c, Ft, Fo, out, st)
        assert len(res) > 100
        # FP32 strains near 1 carry ~1e-7 absolute: e_mu * 1e-7 / q_y relative
        assert np.abs(res).max() < 2e-3, (exp, np.abs(res).max())


def test_pradhana_shift_moves_the_tip_decision_and_resets():
    c = dp_cfg()
    _, K, _, _ = moduli(c)
    p_tip = -c.q_cohesion / c.M
    # pure dilation 100 Pa short of the tip: elastic without a shift
    vol = -(p_tip + 100.0) / K
    F = np.diag(np.full(3, math.exp(vol / 3))).astype(np.float32)[None]
    prad_in = 2 * 100.0 / K                     # shift -K prad = -200 Pa: left of the tip
    st = np.array([[0.0, 0.0, prad_in]], np.float32)
    p0, _, _ = p_q(c, F[0])
    Fo, out, ch, cnt, _ = orc.return_map_state(c, F, st)
    assert ch[0] and cnt[0] == 1
    np.testing.assert_allclose(-K * hencky(Fo[0]).sum(), p_tip, rtol=0, atol=K * 2e-6)   # the original tip
    inst = (p_tip - p0) / K                                                   # < 0: volume given back
    assert inst < 0
    np.testing.assert_allclose(out[0, 2], prad_in + inst, rtol=1e-3)
    np.testing.assert_allclose(out[0, 1], inst, rtol=1e-3)
    # without Pradhana the same state is elastic and nothing accumulates
    c0 = dp_cfg(pradhana=0)
    Fo, out, ch, cnt, _ = orc.return_map_state(c0, F, st)
    assert not ch[0] and cnt[0] == 0
    np.testing.assert_array_equal(Fo[0], F[0])
    assert out[0, 1] == 0.0 and out[0, 2] == np.float32(prad_in)
    # a small accumulated volume and an elastic state right of the shifted tip: reset
    st2 = np.array([[0.0, 0.0, 10.0 / K]], np.float32)
    Fo, out, ch, cnt, _ = orc.return_map_state(c, F, st2)
    assert not ch[0] and cnt[0] == 0 and out[0, 2] == 0.0 and out[0, 3] == 0.0


def test_isotropy():
    c = dp_cfg(visc=0.4)
    r = rng(14)
    for _ in range(40):
        Ft = trial_F(r, vol=r.uniform(-1e-2, 4e-2), dev_amp=r.uniform(0, 1e-2))
        Q = rand_rot(r)
        Fo, *_ = orc.return_map_state(c, Ft[None])
        Fr, *_ = orc.return_map_state(c, (Q @ Ft).astype(np.float32)[None])
        np.testing.assert_allclose(Fr[0], Q @ Fo[0], atol=5e-6)


def test_paint_app_scene_exercises_both_branches():
    """The paint app's material on the oracle: MUSL + DPVisc + Pradhana with the
    spatula on keyframes.  Mass stays put, both the tip and the Perzyna branch
    occur, the accumulator holds the tip's volume gain where it is nonzero."""
    s = scenes.paint_app_2blob()
    o = orc.from_scene(s)
    mp = s["cfg"].rho * s["cfg"].dx ** 3 / s["cfg"].ppc
    o.step(1)
    m0 = o.grid()["mass"].sum(dtype=np.float64)
    np.testing.assert_allclose(m0, s["x"].shape[0] * mp, rtol=1e-5)
    o.step(99)
    out = o.download()
    st = o.stats()
    assert st["exit"] == 0
    assert out["active"].all()
    assert np.isfinite(out["x"]).all() and np.isfinite(out["F"]).all()
    assert st["plastic_count"] > 100
    tip = out["eps_vol"] != 0
    assert tip.sum() > 100, "tension tip reached"
    assert (out["eps"] > 0).sum() > tip.sum(), "Perzyna branch beyond the tip particles"
    # the accumulator holds volume only where the tip was reached, and most
    # tip particles have since moved right of the shifted tip and reset it
    assert np.all(out["eps_vol"][out["prad"] != 0] != 0)
    assert 0 < (out["prad"] != 0).sum() < tip.sum()
    o.step(1)
    np.testing.assert_allclose(o.grid()["mass"].sum(dtype=np.float64), m0, rtol=1e-6)
