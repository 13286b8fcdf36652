"""The C++ Simulation facade driven as the reference's renderer adapter drives
the reference (tests/cpp/adapter_style_demo.cpp, which includes only
matter_compat.hpp): the paint app's scene from colour ratios, latents and the
spatula's animation file (the committed head of the reference's
spatula_motion.bin), two frames on the device, and the public state the
renderer reads."""
import json
import os
import subprocess

import pytest

from pigment_mixing_amd import build as b
from tests.test_gpu_parity import log

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_adapter_style_host_runs_the_paint_scene():
    r = subprocess.run([b.ADAPTER_DEMO, os.path.join(REPO, "pigment_mixing_amd", "config"),
                        os.path.join(REPO, "tests", "golden", "spatula_motion_head.bin"), "2"],
                       capture_output=True, text=True, timeout=240, cwd=REPO)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"[facade] adapter-style: {out}")
    assert out["dim"] == 3 and out["frames"] == 2 and out["exit"] == 0
    assert out["Np"] > 10000 and out["active"] == out["Np"]
    assert 0.02 < out["dx"] < 0.05                      # the sampler's dx at radius 0.01, ppc 27
    assert out["ymin"] >= 0.0                           # on the floor
    assert abs(out["pig4_mean"] - 1.0) < 1e-4           # four concentrations sum to 1
    assert out["inv_row3"] == [0.0, 0.0, 0.0, 1.0]
    assert out["half"] == pytest.approx([0.368, 0.07, 0.736])
    assert out["D_edge"] == pytest.approx([1.5, 7.0])   # pigment_config.json
    assert out["frames_in_file"] == 181
    assert out["lo"][0] <= -2.0 and out["hi"][0] >= 2.0 and out["hi"][1] >= 2.0
