"""CPU checker of the learned dipole guess network (test infrastructure only).

The network has no counterpart in the reference (the reference always starts
the dipole iteration from alpha * E_fixed, MBPolReferenceElectrostaticsForce.cpp:628-639),
so this restates the library's own definition: 48 inputs normalised by
(x - mean) * inv_std, two SiLU layers of 64, a linear layer of 3, loss
(1/n) sum_i |y_i - t_i / scale|^2.  Parameters are one flat vector
W1[48][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3] (matrices [in][out]).

``bf16=True`` rounds, as the MFMA kernels do, the operands of the four matrix
products to bfloat16 (round to nearest even): the normalised inputs, W1, W2,
the first hidden activations, and the two back-propagated pre-activation
gradients.  Everything else is float64 here.  The backward pass is pinned by
finite differences in tests/test_mlp_oracle.py (float64, bf16=False).
"""
import numpy as np

NF, NH, NO = 48, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + NF * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NO
NPARAM = OFF_B3 + NO


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NF, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NH, NO), p[OFF_B3:NPARAM])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _round(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def forward(x, params, mean, inv_std, bf16_ops=True):
    W1, b1, W2, b2, W3, b3 = unpack(params)
    xn = (np.asarray(x, np.float64) - mean) * inv_std
    xb = _round(xn, bf16_ops)
    z1 = xb @ _round(W1, bf16_ops) + b1
    h1 = _round(silu(z1), bf16_ops)
    z2 = h1 @ _round(W2, bf16_ops) + b2
    h2 = silu(z2)
    return h2 @ W3 + b3


def loss_and_grad(x, t, params, mean, inv_std, scale, bf16_ops=True):
    W1, b1, W2, b2, W3, b3 = unpack(params)
    n = len(x)
    xn = (np.asarray(x, np.float64) - mean) * inv_std
    xb = _round(xn, bf16_ops)
    W1b, W2b = _round(W1, bf16_ops), _round(W2, bf16_ops)
    z1 = xb @ W1b + b1
    h1 = _round(silu(z1), bf16_ops)
    z2 = h1 @ W2b + b2
    h2 = silu(z2)
    y = h2 @ W3 + b3
    d = y - np.asarray(t, np.float64) / scale
    loss = float(np.sum(d * d) / n)
    dy = 2.0 * d / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = (h2.T @ dy).ravel()
    g[OFF_B3:NPARAM] = dy.sum(0)
    dz2 = (dy @ W3.T) * dsilu(z2)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz2b = _round(dz2, bf16_ops)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2b).ravel()
    dz1 = (dz2b @ W2b.T) * dsilu(z1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (xb.T @ _round(dz1, bf16_ops)).ravel()
    return loss, g
