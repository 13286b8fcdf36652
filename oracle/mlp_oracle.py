"""CPU oracle of the learned return-map closure's network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
forward / backward kernels, never by the product.

The network is the closure this build adds (the reference has no learned
closure; its exact map is plasticity.cpp:163-238):
8 -> 64 -> 64 -> 64 -> 64 -> 4, SiLU on the hidden layers.  Parameters are a
flat vector, per layer W[out][in] (row-major) then b[out].  With
``bf16=True`` values are rounded to BF16 (round to nearest even) at exactly the
points where the device kernels round: normalised inputs, weights, hidden
activations SiLU(Z + b), the output gradient dY, and each hidden gradient dZ.
Accumulation is FP32.  With ``bf16=False`` and float64 arrays the same
functions give the exact network, which the finite-difference test pins.

Loss: mean over samples and the 4 outputs of the squared normalised error.
"""
import numpy as np

DIMS = (8, 64, 64, 64, 64, 4)


def n_params():
    return sum(DIMS[i + 1] * DIMS[i] + DIMS[i + 1] for i in range(5))


def unpack(p):
    out, k = [], 0
    for i in range(5):
        nin, nout = DIMS[i], DIMS[i + 1]
        W = p[k:k + nout * nin].reshape(nout, nin)
        k += nout * nin
        b = p[k:k + nout]
        k += nout
        out.append((W, b))
    assert k == p.size
    return out


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _silu(z):
    return z / (1 + np.exp(-z))


def _silu_grad(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def _norm(norm24):
    n = np.asarray(norm24)
    return n[0:8], n[8:16], n[16:20], n[20:24]


def forward(params, norm24, x8, bf16_round=True):
    """Network outputs in output units, and the cache for backward."""
    dt = np.float32 if bf16_round else np.float64
    r = bf16 if bf16_round else (lambda a: a)
    im, isd, om, osd = (v.astype(dt) for v in _norm(norm24))
    layers = [(r(W.astype(dt)), b.astype(dt)) for W, b in unpack(np.asarray(params, dtype=dt))]
    a = r(((np.asarray(x8, dtype=dt) - im) / isd).astype(dt))
    acts, zs = [a], []
    for l in range(4):
        W, b = layers[l]
        z = (a @ W.T + b).astype(dt)
        a = r(_silu(z).astype(dt))
        zs.append(z)
        acts.append(a)
    W, b = layers[4]
    yn = (a @ W.T + b).astype(dt)
    return yn * osd + om, dict(layers=layers, acts=acts, zs=zs, yn=yn, r=r, dt=dt)


def loss_and_grad(params, norm24, x8, t4, bf16_round=True):
    """Loss and its gradient (flat, parameter layout) for targets t4 in output units."""
    y, c = forward(params, norm24, x8, bf16_round)
    dt, r = c["dt"], c["r"]
    _, _, om, osd = (v.astype(dt) for v in _norm(norm24))
    tn = ((np.asarray(t4, dtype=dt) - om) / osd).astype(dt)
    n = tn.shape[0]
    e = c["yn"] - tn
    loss = float(np.sum(e.astype(np.float64) ** 2) / (4 * n))
    dY = (2 * e / (4 * n)).astype(dt)
    grads = [None] * 5
    dz_full, dz = dY, r(dY)
    for l in range(4, -1, -1):
        W, _ = c["layers"][l]
        a_in = c["acts"][l]
        grads[l] = ((dz.T @ a_in).astype(dt), dz_full.sum(axis=0).astype(dt))
        if l == 0:
            break
        dA = (dz @ W).astype(dt)
        dz_full = (dA * _silu_grad(c["zs"][l - 1])).astype(dt)
        dz = r(dz_full)
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return loss, flat


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(5):
        nin, nout = DIMS[i], DIMS[i + 1]
        a = np.sqrt(6.0 / (nin + nout))
        parts.append(rng.uniform(-a, a, size=nout * nin))
        parts.append(rng.uniform(-0.1, 0.1, size=nout))
    return np.concatenate(parts).astype(np.float32)
