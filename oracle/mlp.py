"""CPU restatement (float64) of the learned proposal network, its loss and
its gradients.  Test infrastructure: it checks the GPU training and
inference kernels (v3dsfmtoolkit_amd/csrc/k_mlp.hip); nothing in the product
path imports it.

Network (no counterpart in the reference, which matches exhaustively in
V3D/src/CUDA/v3d_cudamatching.cpp:97-127): x = u / 64,
a1 = W1r x + b1, h = gelu(a1) (erf form), z = W2r h + b2,
e = z / sqrt(|z|^2 + 1e-12), with W1r, W2r the BF16 roundings (nearest even)
of the master weights.  Loss over B matched pairs, s_ij = e(l_i).e(r_j)/tau:
L = 1/(2B) sum_i (lse_j s_ij - s_ii) + 1/(2B) sum_j (lse_i s_ij - s_jj).
Gradients are taken with respect to the forward (rounded) weights, which is
what the GPU applies to the masters.

Pinned by tests/test_mlp_oracle.py: central finite differences of `loss`
against `loss_grad` on every parameter block, the BF16 rounding against
torch, and special cases (identical embeddings, tau scaling).
"""
import numpy as np
from scipy.special import erf

IN, HID, OUT = 128, 256, 64
OFF_W1 = 0
OFF_B1 = HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + OUT * HID
N_PARAMS = OFF_B2 + OUT


def bf16_round(x) -> np.ndarray:
    """FP32 -> BF16 -> FP32, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def split(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(HID, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(OUT, HID),
            p[OFF_B2:N_PARAMS])


def forward_params(master):
    """The weights the forward pass uses (float64): rounded W1, W2; biases as is."""
    W1, b1, W2, b2 = split(np.asarray(master, np.float32))
    return (bf16_round(W1).astype(np.float64), b1.astype(np.float64), bf16_round(W2).astype(np.float64),
            b2.astype(np.float64))


def gelu(a):
    return 0.5 * a * (1.0 + erf(a / np.sqrt(2.0)))


def gelu_grad(a):
    return 0.5 * (1.0 + erf(a / np.sqrt(2.0))) + a * np.exp(-0.5 * a * a) / np.sqrt(2.0 * np.pi)


def embed_fp(fp, U):
    """Embeddings of uint8 rows U with forward params fp; returns (e, cache)."""
    W1, b1, W2, b2 = fp
    x = np.asarray(U, np.float64) / 64.0
    a1 = x @ W1.T + b1
    h = gelu(a1)
    z = h @ W2.T + b2
    n = np.sqrt(np.sum(z * z, axis=1, keepdims=True) + 1e-12)
    return z / n, (x, a1, h, z, n)


def embed(master, U):
    return embed_fp(forward_params(master), U)[0]


def _lse(S, axis):
    m = S.max(axis=axis, keepdims=True)
    return (m + np.log(np.exp(S - m).sum(axis=axis, keepdims=True))).squeeze(axis)


def loss_fp(fp, UL, UR, tau):
    B = len(UL)
    e, _ = embed_fp(fp, np.concatenate([UL, UR]))
    S = e[:B] @ e[B:].T / tau
    d = np.diag(S)
    return float((np.sum(_lse(S, 1) - d) + np.sum(_lse(S, 0) - d)) / (2 * B))


def loss_grad_fp(fp, UL, UR, tau):
    """(loss, [dW1, db1, dW2, db2]) with respect to the forward params."""
    W1, b1, W2, b2 = fp
    B = len(UL)
    e, (x, a1, h, z, n) = embed_fp(fp, np.concatenate([UL, UR]))
    EL, ER = e[:B], e[B:]
    S = EL @ ER.T / tau
    lr, lc = _lse(S, 1), _lse(S, 0)
    d = np.diag(S)
    loss = float((np.sum(lr - d) + np.sum(lc - d)) / (2 * B))
    G = (np.exp(S - lr[:, None]) + np.exp(S - lc[None, :]) - 2 * np.eye(B)) / (2 * B)
    dE = np.concatenate([G @ ER, G.T @ EL]) / tau
    dz = (dE - e * np.sum(e * dE, axis=1, keepdims=True)) / n
    dW2 = dz.T @ h
    db2 = dz.sum(0)
    da1 = (dz @ W2) * gelu_grad(a1)
    dW1 = da1.T @ x
    db1 = da1.sum(0)
    return loss, [dW1, db1, dW2, db2]


def loss_grad(master, UL, UR, tau):
    """(loss, flat gradient in the parameter layout W1 | b1 | W2 | b2)."""
    loss, g = loss_grad_fp(forward_params(master), UL, UR, tau)
    return loss, np.concatenate([x.reshape(-1) for x in g])


def topk(EQ, ER, k=16):
    """Top-k right indices per query by embedding similarity (descending,
    ties towards the smaller index)."""
    S = np.asarray(EQ, np.float64) @ np.asarray(ER, np.float64).T
    return np.argsort(-S, axis=1, kind="stable")[:, :k]
