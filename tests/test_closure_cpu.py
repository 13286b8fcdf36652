"""CPU checks of the learned closure's network oracle (oracle/mlp_oracle.py)
and of the library's parameter layout.

The oracle's backward pass is pinned by central finite differences on the
exact (float64, unrounded) network; its BF16 rounding by known answers."""
import os
import sys

import numpy as np

from pigment_mixing_amd import abi

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def _norm24(rng):
    return np.concatenate([rng.normal(0, 0.5, 8), rng.uniform(0.5, 2.0, 8),
                           rng.normal(0, 0.5, 4), rng.uniform(0.5, 2.0, 4)])


def test_parameter_count_matches_library():
    assert abi.load().mpmx_closure_n_params() == mo.n_params() == 13316


def test_bf16_rounding_known_answers():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)           # BF16 spacing at 1.0
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20,
                     -2.5, 3.0e-3, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -2.5, np.float32(0.0030059814453125), 0.0],
                    np.float32)
    # ties go to even: 1 + half-ulp -> 1; 1 + 1.5 ulp -> 1 + 2 ulp; above the tie rounds up
    np.testing.assert_array_equal(mo.bf16(vals), want)
    assert mo.bf16(np.array([one]))[0] == one


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = mo.init_params(5).astype(np.float64)
    n24 = _norm24(rng)
    x = rng.normal(0, 1, (37, 8))
    t = rng.normal(0, 1, (37, 4))
    loss, g = mo.loss_and_grad(p, n24, x, t, bf16_round=False)
    assert loss > 0
    idx = rng.choice(p.size, 300, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, n24, x, t, False)[0] - mo.loss_and_grad(pm, n24, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_rounded_gradient_close_to_exact():
    """The BF16-rounded oracle stays within BF16 precision of the exact gradient
    (the rounding points add relative errors of order 2^-8)."""
    rng = np.random.default_rng(4)
    p = mo.init_params(6)
    n24 = _norm24(rng)
    x = rng.normal(0, 1, (256, 8)).astype(np.float32)
    t = rng.normal(0, 1, (256, 4)).astype(np.float32)
    _, g64 = mo.loss_and_grad(p.astype(np.float64), n24, x.astype(np.float64), t.astype(np.float64), False)
    _, g16 = mo.loss_and_grad(p, n24, x, t, True)
    assert np.linalg.norm(g16 - g64) <= 3e-2 * np.linalg.norm(g64)
