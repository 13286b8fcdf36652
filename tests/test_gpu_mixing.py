"""Pigment mixing on the device: the closed-form check of
tests/test_pigment_mixing.py on the HIP step, and a knife scene in which the
mixing is active, against the oracle after 100 steps."""
import numpy as np
import pytest

from pigment_mixing_amd import abi, scenes
from tests import oracle_py as orc
from tests.test_gpu_parity import check_particles, log, run_both, small
from tests.test_pigment_mixing import check_pigment_mixing

pytestmark = pytest.mark.gpu


def test_pigment_mixing_closed_form_on_the_device():
    check_pigment_mixing(lambda cfg: abi.Handle(cfg), log=log)


def test_parity_knife_scene_with_active_mixing_100_steps():
    """The small knife scene with the mixing edges lowered (D_edge0 0.05,
    D_edge1 2 s^-1, so that the knife's shear mixes the two latents; the
    reference's 1.5 / 7 s^-1 leave them untouched at this scale) and the boost
    window open: GPU and oracle agree at the parity bars and the latents moved."""
    s = small()
    c = s["cfg"]
    c.D_edge0, c.D_edge1 = 0.05, 2.0
    c.start_boost_time, c.end_boost_time, c.boost_factor = 0.0, 5e-4, 3.0
    h, o = run_both(s, 100)
    g, r = h.download(), o.download()
    moved = np.abs(r["pig"] - s["pig"]).max()
    log(f"[mixing] knife scene: max |pig - pig0| (oracle) = {moved:.2e}")
    assert moved > 1e-4
    check_particles(g, r, scenes.DX, "small/knife/mixing/100")
