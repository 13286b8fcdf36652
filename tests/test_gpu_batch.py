"""GPU checks of batched scenes (one handle, several independent scenes in the
same kernels and graph) and of the launch-shape autotuner.

A scene inside a batch must behave as the same scene alone: the only
difference is FP32 atomic order in the scatter, so the small-scene parity bars
apply (x <= 1e-4 dx, v <= 1e-3 relative with a 0.1 m/s floor, F <= 1e-4
relative, latents <= 1e-5).  One batched scene is also checked against the
CPU oracle."""
import numpy as np
import pytest

from pigment_mixing_amd import abi, scenes
from tests import oracle_py as orc
from tests.test_gpu_parity import check_particles, log

pytestmark = pytest.mark.gpu


def small_batch(n=3):
    return [scenes.batch_scene(s, seed=100, np_range=(3000, 9000)) for s in range(n)]


def split(out, rng):
    a, n = rng
    return {k: v[a:a + n] for k, v in out.items()}


def test_batched_scenes_match_individual_runs():
    sl = small_batch(3)
    log("[batch] scenes " + ", ".join(f"{s['name']} n={s['x'].shape[0]} yield={s['yield_pa']:.1f}" for s in sl))
    hb = scenes.make_batch_handle(sl)
    hb.step(60)
    ob = hb.download()
    for i, sc in enumerate(sl):
        h1 = scenes.make_handle(sc)
        h1.step(60)
        check_particles(split(ob, hb.scene_range(i)), h1.download(), scenes.DX, f"batch/{i}", v_tol=1e-3)


def test_batched_scene_matches_oracle():
    sl = small_batch(3)
    hb = scenes.make_batch_handle(sl)
    hb.step(40)
    ob = hb.download()
    o = orc.from_scene(sl[1])
    o.step(40)
    check_particles(split(ob, hb.scene_range(1)), o.download(), scenes.DX, "batch/1 vs oracle", v_tol=1e-3)


def test_autotune_keeps_results():
    sl = small_batch(3)
    ha = scenes.make_batch_handle(sl)
    hb = scenes.make_batch_handle(sl)
    r = ha.autotune(2)
    log(f"[batch] autotune: p2g {r.wg_p2g}/CU ({r.ms_p2g:.4f} ms) g2p {r.wg_g2p}/CU ({r.ms_g2p:.4f} ms) "
        f"grid {r.wg_grid}/CU ({r.ms_grid:.4f} ms), {r.trials} trials, {r.steps} steps")
    assert r.wg_p2g in (8, 12, 16, 20, 24, 32) and r.wg_g2p in (2, 3, 4, 6, 8) and r.wg_grid in (2, 4, 8)
    assert r.steps == r.trials * 2
    ha.step(30)
    hb.step(30 + r.steps)
    assert ha.stats().step == hb.stats().step
    check_particles(ha.download(), hb.download(), scenes.DX, "autotuned vs default", v_tol=1e-3)
