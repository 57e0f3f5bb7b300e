"""FP64 reference of the warm-start network (test infrastructure).

Forward pass, mean-squared-error loss and back-propagated gradients of the
22 -> 128 -> 128 -> 128 -> 3 GELU MLP whose parameters are packed as
fc3d_hip.h describes (W1[128][32], b1, W2[128][128], b2, W3[128][128], b3,
W4[16][128], b4[16]).

``emulate=True`` rounds to BF16 where the device kernels do:
- the standardised inputs;
- the weights;
- the hidden activations;
- the back-propagated dZ.

Everything else stays in FP64. The device accumulates in FP32, so results
agree to accumulation-order level, not bit for bit.

``mutation`` plants a mistake for the tests that show the finite-difference
check catches it: "gelu_grad" drops the second term of GELU's derivative.
"""
import numpy as np
from scipy.special import erf

NF, K1, H, NO, NOP = 22, 32, 128, 3, 16
SHAPES = [(H, K1), (H, H), (H, H), (NOP, H)]
NPARAMS = sum(n * k + n for n, k in SHAPES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    off, Ws, bs = 0, [], []
    for n, k in SHAPES:
        Ws.append(p[off:off + n * k].reshape(n, k))
        off += n * k
        bs.append(p[off:off + n])
        off += n
    assert off == len(p) == NPARAMS
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in zip(Ws, bs)])


def bf16(x):
    """Round (via float32) to the nearest BF16, ties to even; as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z, mutation=None):
    g = 0.5 * (1.0 + erf(z / np.sqrt(2.0)))
    if mutation != "gelu_grad":
        g = g + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)
    return g


def standardise(x, mean, std, emulate):
    if emulate:
        xs = (np.asarray(x, np.float32) - np.asarray(mean, np.float32)) / np.asarray(std, np.float32)
        xs = bf16(xs)
    else:
        xs = (np.asarray(x, np.float64) - mean) / std
    a = np.zeros((len(xs), K1))
    a[:, :NF] = xs
    return a


def forward(p, x, mean, std, emulate=False):
    Ws, bs = unpack(p)
    if emulate:
        Ws = [bf16(W) for W in Ws]
    a = standardise(x, mean, std, emulate)
    acts, zs = [a], []
    for layer in range(3):
        z = a @ Ws[layer].T + bs[layer]
        zs.append(z)
        a = gelu(z)
        if emulate:
            a = bf16(a)
        acts.append(a)
    y = a @ Ws[3].T + bs[3]
    return y[:, :NO], (Ws, zs, acts)


def loss_and_grads(p, x, t, mean, std, emulate=False, mutation=None):
    n = len(x)
    y, (Ws, zs, acts) = forward(p, x, mean, std, emulate)
    diff = y - np.asarray(t, np.float64)
    loss = float((diff ** 2).sum() / (n * NO))
    dz = np.zeros((n, NOP))
    dz[:, :NO] = 2.0 * diff / (n * NO)
    if emulate:
        dz = bf16(dz)
    gW = [None] * 4
    gb = [None] * 4
    gW[3] = dz.T @ acts[3]
    gb[3] = dz.sum(axis=0)
    da = dz @ Ws[3]
    for layer in (2, 1, 0):
        dz = da * gelu_grad(zs[layer], mutation)
        if emulate:
            dz = bf16(dz)
        gW[layer] = dz.T @ acts[layer]
        gb[layer] = dz.sum(axis=0)
        if layer > 0:
            da = dz @ Ws[layer]
    return loss, pack(gW, gb)


def block_slices():
    """(name, slice) of every weight and bias block in the packed vector."""
    out, off = [], 0
    for i, (n, k) in enumerate(SHAPES):
        out.append((f"W{i + 1}", slice(off, off + n * k)))
        off += n * k
        out.append((f"b{i + 1}", slice(off, off + n)))
        off += n
    return out
