"""GPU checks of batched scenes (one handle, several independent scenes in the
same kernels and graph) and of the launch-shape autotuner.

A scene inside a batch must behave bit for bit as the same scene alone: the
sort orders the particles of a cell by position (k_cellsort), so the
scatter's FP32 sums run in the same order whatever the other scenes of the
batch, the atomic timing or the launch shapes are.  One batched scene is also
checked against the CPU oracle at the parity bars (x <= 1e-4 dx, v <= 1e-4
relative with a 0.1 m/s floor, F <= 1e-4 relative, latents <= 1e-5)."""
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


def assert_identical(a, b, label):
    """Every downloaded field equal bit for bit (NaN-free state)."""
    for k in ("x", "v", "F", "C", "pig", "diff", "eps", "dgamma", "active", "eps_vol", "prad"):
        diff = np.count_nonzero(a[k] != b[k])
        assert diff == 0, f"{label}: {diff} values of {k} differ"
    log(f"[batch] {label}: bit-identical ({len(a['x'])} particles)")


def test_batched_scenes_match_individual_runs():
    sl = small_batch(3)
    log("[batch] scenes " + ", ".join(f"{s['name']} n={s['x'].shape[0]} yield={s['yield_pa']:.1f}" for s in sl))
    hb = scenes.make_batch_handle(sl)
    hb.step(60)
    ob = hb.download()
    for i, sc in enumerate(sl):
        h1 = scenes.make_handle(sc)
        h1.step(60)
        assert_identical(split(ob, hb.scene_range(i)), h1.download(), f"batch/{i} vs alone")


def test_two_runs_are_bit_identical():
    """Reproducibility: the same batch stepped twice in two handles (USL and
    MUSL) downloads bit-identical state, and so does the grid."""
    for musl in (0, 1):
        sl = small_batch(3)
        for sc in sl:
            sc["cfg"].use_musl = musl
        ha = scenes.make_batch_handle(sl)
        hb = scenes.make_batch_handle(sl)
        ha.step(50)
        hb.step(50)
        assert_identical(ha.download(), hb.download(), f"two runs (musl={musl})")
        ga, gb = ha.grid(), hb.grid()
        for k in ga:
            assert np.array_equal(ga[k], gb[k]), f"grid {k} differs between runs"


def test_batched_scene_matches_oracle():
    sl = small_batch(3)
    hb = scenes.make_batch_handle(sl)
    hb.step(40)
    ob = hb.download()
    o = orc.from_scene(sl[1])
    o.step(40)
    check_particles(split(ob, hb.scene_range(1)), o.download(), scenes.DX, "batch/1 vs oracle")


def test_autotune_keeps_results():
    sl = small_batch(3)
    ha = scenes.make_batch_handle(sl)
    hb = scenes.make_batch_handle(sl)
    r = ha.autotune(2)
    log(f"[batch] autotune: p2g {r.wg_p2g}/CU ({r.ms_p2g:.4f} ms) g2p {r.wg_g2p}/CU ({r.ms_g2p:.4f} ms) "
        f"grid {r.wg_grid}/CU ({r.ms_grid:.4f} ms), {r.trials} trials, {r.steps} steps")
    # P2G is tuned up to its occupancy (at most 16 one-wavefront workgroups per CU here)
    assert 1 <= r.wg_p2g <= 16 and r.wg_g2p in (2, 3, 4, 6, 8, 12, 16) and r.wg_grid in (2, 4, 8)
    assert r.steps == r.trials * 2
    ha.step(30)
    hb.step(30 + r.steps)
    assert ha.stats().step == hb.stats().step
    assert_identical(ha.download(), hb.download(), "autotuned vs default shapes")


@pytest.mark.parametrize("musl", [0, 1])
def test_launch_shapes_keep_results(musl):
    """P2G items and G2P blocks go to wavefronts first statically (four each)
    and then through per-XCD queues; which wavefront processes what depends on
    the launch shape and on timing, the results must not.  A 590k-particle
    scene (~1,660 particle blocks) at one workgroup per CU (32 walkers per
    XCD: ~6 G2P blocks and ~26 P2G items each) runs mostly from the queues;
    shapes beyond the kernels' occupancy start some static shares late.
    Every shape, set with mpmx_set_launch_shape, must give the default shape's
    state bit for bit over steps that reset and reuse the queues in every
    launch (USL, and MUSL with its second scatter)."""
    sc = scenes.paint_4blob(cells=(48, 8, 48), dims=(192, 64, 192))
    sc["cfg"].use_musl = musl
    ref = scenes.make_handle(sc)
    ref.step(12)
    want = ref.download()
    for shape in ((1, 1, 1), (24, 16, 8), (5, 3, 2)):
        h = scenes.make_handle(sc)
        h.set_launch_shape(*shape)
        h.step(12)
        assert h.stats().device_error == 0
        assert_identical(h.download(), want, f"launch shape {shape} (musl={musl})")
        h.close()
    with pytest.raises(RuntimeError):
        ref.set_launch_shape(0, 4, 4)


def test_batch_mixing_drucker_prager_and_von_mises_scenes():
    """A batch holding the paint app's Drucker-Prager scene (volume rows,
    Pradhana accumulation) next to a von Mises scene (both MUSL, as a batch
    shares the MUSL flag): each equals its own run.  The Drucker-Prager scene
    starts from the oracle's step-25 state (1,599 particles at the tension tip,
    restarted through mpmx_upload_plastic_state on an explicit lattice origin),
    so that the comparison runs where the map is active.  Both scenes must
    match their own runs bit for bit: the Drucker-Prager scene runs the same
    (volume-tracking) G2P variant in both handles, the von Mises scene the
    volume-tracking variant in the batch and the plain one alone (the same
    arithmetic in two instantiations), and it leaves the volume rows at zero."""
    app = scenes.paint_app_2blob()
    o0 = orc.from_scene(app)
    o0.step(25)
    state = o0.download()
    _, origin, *_ = o0.lattice()
    cfg = abi.Config.from_buffer_copy(app["cfg"])
    cfg.origin_mode = 1
    cfg.grid_origin[:] = [float(v) for v in origin]
    vm = scenes.batch_scene(1, seed=100, np_range=(3000, 9000))
    vm["cfg"].use_musl = 1
    args = dict(v=state["v"], F=state["F"], C=state["C"], pig=state["pig"], diff=state["diff"], eps=state["eps"],
                active=state["active"])
    hb = abi.Handle([cfg, vm["cfg"]])
    hb.upload_scene(0, state["x"], **args)
    hb.upload_plastic_state(state["eps_vol"], state["prad"], scene=0)
    hb.set_objects_scene(0, app["plates"], None)
    hb.upload_scene(1, vm["x"], pig=vm["pig"])
    hb.set_objects_scene(1, vm["plates"], vm["knife"])
    hb.prepare()
    hb.step(5)
    ob = hb.download()
    h0 = abi.Handle(cfg)
    h0.upload(state["x"], **args)
    h0.upload_plastic_state(state["eps_vol"], state["prad"])
    h0.set_objects(app["plates"], None)
    h0.prepare()
    h1 = scenes.make_handle(vm)
    for i, h in enumerate((h0, h1)):
        h.step(5)
        a, b = split(ob, hb.scene_range(i)), h.download()
        np.testing.assert_array_equal(a["active"], b["active"])
        if i == 1:
            assert not a["eps_vol"].any() and not a["prad"].any()
        assert_identical(a, b, f"mixed/{i} ({'Drucker-Prager' if i == 0 else 'von Mises'}) vs alone")


def test_batch_64_scenes_full_workload():
    """The batched workload at its own size: 64 scenes of 20k-400k particles on
    64^3-160^3 lattices (uneven rheology and knife speed), one handle.  After 40
    steps the smallest scene matches the CPU oracle at the parity bars; after 100
    steps no device error is set, every particle is live and inside its own
    scene's lattice, scene 0's node mass equals its particles' mass (the
    float-coordinate partition of unity, as the large-scene test), and the
    smallest and largest scenes equal their own single-scene runs bit for bit."""
    sl = scenes.batch_scenes(64)
    sizes = np.array([sc["x"].shape[0] for sc in sl])
    lo, hi = int(sizes.argmin()), int(sizes.argmax())
    log(f"[batch64] {sizes.sum()} particles, sizes {sizes.min()}..{sizes.max()}, lattices "
        f"{min(sc['cfg'].grid_dims[0] for sc in sl)}..{max(sc['cfg'].grid_dims[0] for sc in sl)}")
    assert sizes.min() >= 15_000 and sizes.max() <= 450_000
    hb = scenes.make_batch_handle(sl)
    hb.step(40)
    ob = hb.download()
    o = orc.from_scene(sl[lo])
    o.step(40)
    check_particles(split(ob, hb.scene_range(lo)), o.download(), scenes.DX, f"batch64/{lo} vs oracle")
    del o, ob
    m0 = hb.grid()["mass"].sum(dtype=np.float64)
    hb.step(60)
    st = hb.stats()
    assert st.device_error == 0 and st.stray_count == 0
    assert st.n_active == sizes.sum(), f"{sizes.sum() - st.n_active} particles deactivated"
    ob = hb.download()
    assert np.isfinite(ob["x"]).all() and np.isfinite(ob["v"]).all() and np.isfinite(ob["F"]).all()
    for i, sc in enumerate(sl):
        a, n = hb.scene_range(i)
        x = ob["x"][a:a + n].astype(np.float64)
        org = sc["x"].min(0).astype(np.float64) - 2 * sc["cfg"].dx     # the lattice origin rule (bbox_min - 2dx)
        ext = (np.asarray(sc["cfg"].grid_dims[:], np.float64) - 1) * sc["cfg"].dx
        assert (x >= org).all() and (x <= org + ext).all(), f"scene {i} left its lattice"
    m1 = hb.grid()["mass"].sum(dtype=np.float64)
    np.testing.assert_allclose(m1, m0, rtol=1e-6)
    mp = sl[0]["cfg"].rho * sl[0]["cfg"].dx ** 3 / sl[0]["cfg"].ppc
    np.testing.assert_allclose(m0, sizes[0] * mp, rtol=2e-5)
    for i in (lo, hi):
        h1 = scenes.make_handle(sl[i])
        h1.step(100)
        assert_identical(split(ob, hb.scene_range(i)), h1.download(), f"batch64/{i} vs alone")
        h1.close()
    hb.close()
