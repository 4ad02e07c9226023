"""Host-side driver of the learned proposal path: training pairs labelled by
the exact int8 matcher on synthetic image pairs, and a training schedule for
EmbeddingNet.  The GPU does the labelling, the training and the matching;
this module only orders the calls."""
from __future__ import annotations

import math

import numpy as np

from ._lib import KERNEL_I8, synth_pair, synth_seed
from .matching import EmbeddingNet, SIFTMatching

TRAIN_CONFIG = 40  # generator seed space of training pairs (disjoint from evaluation pairs)


def labeled_pairs(count: int, n: int = 8192, rho: float = 0.4, sigma: float = 0.05, min_score: float = 0.8,
                  min_ratio: float = 0.8, config: int = TRAIN_CONFIG, device: int = 0):
    """`count` matched descriptor pairs (left rows, right rows), uint8, taken
    from the exact matcher's lists on synthetic n x n image pairs."""
    m = SIFTMatching(device)
    m.enableRoundingWarning(False)
    m.allocate(n, n)
    m.setKernel(KERNEL_I8)
    ls, rs, total, i = [], [], 0, 0
    while total < count:
        d = synth_pair(synth_seed(config, i), n, n, rho, sigma)
        m.setLeftFeatures(n, d["L"])
        m.setRightFeatures(n, d["R"])
        p = np.array(m.findPutativeMatches(min_score, min_ratio), np.int64).reshape(-1, 2)
        ls.append(d["L"][p[:, 0]])
        rs.append(d["R"][p[:, 1]])
        total += len(p)
        i += 1
    m.close()
    return np.concatenate(ls)[:count], np.concatenate(rs)[:count]


def train(net: EmbeddingNet, left: np.ndarray, right: np.ndarray, steps: int, batch: int = 1024,
          lr: float = 2e-3, temperature: float = 0.05, seed: int = 7, segments: int = 8) -> np.ndarray:
    """Adam with a cosine learning-rate decay over `segments` constant pieces;
    returns the per-step batch losses."""
    net.set_training_pairs(left, right)
    losses = []
    per = max(1, steps // segments)
    done = 0
    while done < steps:
        s = min(per, steps - done)
        frac = (done + 0.5 * s) / steps
        losses.append(net.train(s, batch, lr * 0.5 * (1 + math.cos(math.pi * frac)), temperature, seed + done))
        done += s
    return np.concatenate(losses)


def projection_params(seed: int = 3) -> np.ndarray:
    """Network parameters that make EmbeddingNet a random orthogonal 128 -> 64
    projection followed by the L2 normalisation: the same pipeline without
    what training learns (its ablation).  With P (64 x 128, orthonormal rows),
    W1 = [P; -P; 0], b1 = 0, W2 = [I, -I, 0], b2 = 0, the hidden layer gives
    GELU(Px) - GELU(-Px) = Px exactly (GELU(a) - GELU(-a) = a), so the
    embedding is P x / |P x| up to the BF16 rounding of P."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((128, 64)))
    P = q.T.astype(np.float32)  # 64 x 128, orthonormal rows
    W1 = np.zeros((256, 128), np.float32)
    W1[:64] = P
    W1[64:128] = -P
    W2 = np.zeros((64, 256), np.float32)
    W2[:, :64] = np.eye(64, dtype=np.float32)
    W2[:, 64:128] = -np.eye(64, dtype=np.float32)
    return np.concatenate([W1.ravel(), np.zeros(256, np.float32), W2.ravel(), np.zeros(64, np.float32)])
