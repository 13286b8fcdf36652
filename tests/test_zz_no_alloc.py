"""The allocation-free forward pass of the closure network
(mpmx_mlp_forward_noalloc) against the allocating entry point
(mpmx_mlp_forward): the same kernels on the same inputs, so the outputs must be
equal bit for bit.  Two calls in a row, the second smaller and with other
parameters, check that nothing of the first call's contents in the shared
device array leaks into the second.  Sizes beyond the array are refused before
any device work."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from pigment_mixing_amd import abi

pytestmark = pytest.mark.gpu


def _case(seed, n):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    n24 = np.concatenate([rng.normal(0, 0.5, 8), rng.uniform(0.5, 2.0, 8),
                          rng.normal(0, 0.5, 4), rng.uniform(0.5, 2.0, 4)]).astype(np.float32)
    x = (rng.normal(0, 1, (n, 8)) * n24[8:16] + n24[0:8]).astype(np.float32)
    return p, n24, x


def test_noalloc_forward_equals_allocating_forward_twice():
    for seed, n in ((31, 5000), (32, 777)):
        p, n24, x = _case(seed, n)
        print(f"[noalloc] forward seed={seed} n={n}", flush=True)
        y_na = abi.mlp_forward(p, n24, x, noalloc=True)
        y = abi.mlp_forward(p, n24, x)
        assert np.isfinite(y).all()
        bad = np.count_nonzero(y_na != y)
        print(f"[noalloc] seed={seed} n={n}: {bad} of {y.size} outputs differ", flush=True)
        assert bad == 0
        # and the network agrees with its CPU oracle at the forward test's bar
        yo, _ = mo.forward(p, n24, x, True)
        scale = n24[20:24] * max(1e-3, float(np.std((yo - n24[16:20]) / n24[20:24])))
        assert (np.abs(y_na - yo) / scale).max() <= 1e-2


def test_noalloc_forward_refuses_sizes_beyond_its_array():
    p, n24, x = _case(33, 8)
    big = np.zeros((abi.MLP_NOALLOC_MAX + 1, 8), np.float32)
    with pytest.raises(abi.MpmxError) as e:
        abi.mlp_forward(p, n24, big, noalloc=True)
    assert e.value.code == 5      # MPMX_ERR_CAPACITY
