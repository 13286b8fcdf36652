"""The reference's physics tests (matter/tests/tests.cpp) on the oracle, in
short versions that finish in seconds on the CPU (scenes and analytic checks in
tests/reference_scenes.py).  Over these short times the elastic settling of
the block is a visible share of the motion, so the bars are wider than the
reference's (which it applies after 5 s or 100 s); tests/test_gpu_reference_ports.py
runs the long versions on the device with the reference's own bars."""
import pytest

from pigment_mixing_amd import abi
from tests import oracle_py as orc
from tests import reference_scenes as rs


def make(cfg):
    return orc.Oracle(cfg)


@pytest.mark.parametrize("ground", ["plate", "general"])
@pytest.mark.parametrize("bc", [abi.BC_SLIPFREE, abi.BC_SLIPSTICK])
def test_coulomb_friction_incline_oracle(ground, bc):
    """CoulombFrictionTest.{Plate,General}{SlipFree,SlipStick} (tests.cpp:100-306),
    0.2 s: measured 4.8 % (SlipFree: friction only while approaching, so the
    block's bouncing slides freely), 0.1 % (plate SlipStick), 2.3 % (general
    SlipStick)."""
    sc = rs.incline_scene(ground, bc, t_end=0.2)
    _, out = rs.run(sc, make)
    got, ref, err = rs.incline_error(sc, out)
    assert out["active"].all()
    assert err <= (0.06 if bc == abi.BC_SLIPFREE else 0.03), (got, ref, err)


def test_analytic_slide_oracle():
    """BoundaryTest.AnalyticSlipFree (tests.cpp:13-55) on a 30 deg slab: |v| =
    sqrt(2 g h) after 0.2 s within 5e-3 (measured 3.7e-3; the initial settling
    into the slab is the remainder, 1.2e-2 at 0.1 s)."""
    sc = rs.slide_scene(abi.BC_SLIPFREE, t_end=0.2)
    _, out = rs.run(sc, make)
    v, ref, err, drop = rs.slide_error(sc, out)
    assert out["active"].all() and drop > 0.04
    assert err <= 5e-3, (v, ref, err)


def test_bulk_modulus_oracle():
    """ElasticityTest.BulkModulus (tests.cpp:475-540): a 10x10x4-cell block
    squeezed by moving NoSlip plates for 50 ms; K within 3 % (measured 0.2 %)."""
    sc = rs.bulk_scene((10, 10, 4), 500)
    _, out = rs.run(sc, make)
    K, K_true, err, J_avg = rs.bulk_modulus(sc["cfg"], out["F"])
    assert J_avg < 0.9999          # compressed
    assert err <= 0.03, (K, K_true)


def test_rotation_energy_oracle():
    """EnergyTest.Rotation (tests.cpp:542-602): kinetic energy after 30 ms of a
    spinning, translating block: no gain, loss <= 1e-3 (measured 4e-6)."""
    sc = rs.rotation_scene((10, 10, 3), 300)
    _, out = rs.run(sc, make)
    e0, e1 = rs.kinetic(sc["v"]), rs.kinetic(out["v"])
    assert out["active"].all()
    assert 0.0 <= (e0 - e1) / e0 <= 1e-3
