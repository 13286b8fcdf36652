"""GPU checks of the adaptive time step (Simulation::updateDt,
update_dt.cpp:5-71) offered as an option: mpmx_max_speed reduces the live
particles' speed on the device, and mpmx_set_dt changes the step of a running
handle, its knife clock and plates continuing from the current step."""
import ctypes

import numpy as np
import pytest

from pigment_mixing_amd import abi, scenes
from tests import oracle_py as orc
from tests.test_gpu_parity import check_particles, log, small

pytestmark = pytest.mark.gpu


def expected_max_speed(out):
    a = out["active"].astype(bool)
    v = out["v"][a].astype(np.float32)
    v2 = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
    return float(np.sqrt(np.float32(v2.max())))


def test_max_speed_matches_the_particles():
    s = small()
    h = scenes.make_handle(s)
    assert h.max_speed() == 0.0
    h.step(20)
    out = h.download()
    vmax = h.max_speed()
    log(f"[dt] max speed after 20 steps {vmax:.6e}")
    assert vmax > 0.0
    assert vmax == pytest.approx(expected_max_speed(out), rel=1e-6)


def test_adaptive_dt_matches_oracle():
    """The reference's rule each step: dt = min(cfl dx / max|v|, dt_max),
    clamped to the end of the frame and to min_dt (update_dt.cpp:7-41), here
    with dt_max = 3e-5 s and 1,000 frames per second so that the sequence
    holds full and partial steps.  The device and the oracle step with the
    same dt sequence and agree at the parity bars after 45 steps."""
    s = small(yield_pa=5.0, knife=False)
    s["cfg"].fps = 1000.0
    h = scenes.make_handle(s)
    o = orc.from_scene(s)
    cfl, dt_max, min_dt = 0.5, np.float32(3e-5), np.float32(1e-14)
    frame_dt = np.float32(1.0) / np.float32(s["cfg"].fps)
    time, frame = np.float32(0.0), 0
    dts = []
    for step in range(45):
        vmax = h.max_speed()
        dt = np.float32(min(cfl * s["cfg"].dx / vmax, dt_max)) if vmax > 1e-10 else dt_max
        dt = min(dt, np.float32(frame_dt * np.float32(frame + 1) - time))
        dt = max(dt, min_dt)
        h.set_dt(dt)
        o.set_dt(dt)
        h.step(1)
        o.step(1)
        time = np.float32(time + dt)
        if frame_dt * np.float32(frame + 1) - time < min_dt * np.float32(1.1):
            frame += 1
        dts.append(float(dt))
    log(f"[dt] steps {len(dts)}: dt from {min(dts):.3e} to {max(dts):.3e}, {frame} frame ends crossed")
    assert frame >= 1 and min(dts) < 0.9 * float(dt_max), "the sequence must hold a partial step"
    check_particles(h.download(), o.download(), scenes.DX, "adaptive dt/45")


def test_set_dt_continues_the_knife_clock():
    """After a dt change at step 10 (dt -> 2 dt) the knife is evaluated at
    10 dt + (n - 10) 2 dt, i.e. at step n - 5 of a 2 dt clock."""
    s = small()
    dt1 = np.float32(s["cfg"].dt)
    dt2 = np.float32(2.0) * dt1
    h = scenes.make_handle(s)
    h.step(10)
    h.set_dt(dt2)
    h.step(6)
    lib = abi.load()
    for n in (11, 12, 16, 20):
        got = h.knife_state(n)
        ref = abi.KnifeState()
        assert lib.mpmx_knife_state_eval(ctypes.byref(s["knife"]), None, 0, 60.0, float(dt2), n - 5,
                                         ctypes.byref(ref)) == 0
        for f in ("T", "v", "w"):
            np.testing.assert_allclose(np.array(getattr(got, f)), np.array(getattr(ref, f)), rtol=1e-6, atol=1e-7,
                                       err_msg=f"step {n} {f}")
    st = h.stats()
    assert st.step == 16 and st.device_error == 0
