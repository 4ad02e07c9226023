"""Pins the CPU oracles of the learned proposal path: the network's loss
gradient (central finite differences, closed-form loss), and the learned
decision pipeline (reduces to the exact rule when nothing is left to chance)."""
import numpy as np
import pytest
import torch

import v3dsfmtoolkit_amd as v
from oracle import learned as OL
from oracle import mlp as OM
from oracle import oracle as O


def small_fp(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / 128), (256, 128))
    b1 = rng.normal(0, 0.1, 256)
    W2 = rng.normal(0, np.sqrt(1 / 256), (64, 256))
    b2 = rng.normal(0, 0.1, 64)
    return [W1, b1, W2, b2]


def pair_batch(seed, B):
    d = v.synth_pair(v.synth_seed(5, seed), 4 * B, 4 * B, 1.0, 0.05)
    src = d["src"]
    j = np.flatnonzero(src >= 0)[:B]
    return d["L"][src[j]], d["R"][j]


def test_bf16_round_matches_torch():
    x = np.random.default_rng(1).standard_normal(50000).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(OM.bf16_round(x), want)


@pytest.mark.parametrize("block", [0, 1, 2, 3])
def test_gradient_matches_finite_differences(block):
    fp = small_fp(3)
    UL, UR = pair_batch(1, 8)
    tau = 0.1
    _, g = OM.loss_grad_fp(fp, UL, UR, tau)
    rng = np.random.default_rng(block)
    shape = fp[block].shape
    for _ in range(12):
        idx = tuple(int(rng.integers(0, s)) for s in shape)
        h = 1e-5
        fp_p = [x.copy() for x in fp]
        fp_m = [x.copy() for x in fp]
        fp_p[block][idx] += h
        fp_m[block][idx] -= h
        fd = (OM.loss_fp(fp_p, UL, UR, tau) - OM.loss_fp(fp_m, UL, UR, tau)) / (2 * h)
        assert abs(fd - g[block][idx]) <= 1e-6 + 1e-4 * abs(fd), (block, idx, fd, g[block][idx])


def test_loss_closed_form_for_identical_embeddings():
    """W2 = 0: every embedding is b2 / |b2|, all logits equal, loss = log B."""
    fp = small_fp(4)
    fp[2] = np.zeros_like(fp[2])
    UL, UR = pair_batch(2, 16)
    loss, g = OM.loss_grad_fp(fp, UL, UR, 0.05)
    assert loss == pytest.approx(np.log(16), rel=1e-12)
    assert np.allclose(g[3], 0, atol=1e-9)  # moving b2 moves every embedding alike


def test_learned_oracle_reduces_to_exact_rule():
    d = v.synth_pair(v.synth_seed(5, 9), 640, 512, 0.4, 0.05)
    L, R = d["L"], d["R"]
    rng = np.random.default_rng(0)
    for rule in [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)]:
        want = O.find_u8(L, R, *rule)[2]
        # every right a candidate: nothing is left to the proposal
        full = np.tile(np.arange(512), (640, 1))
        tab, _, _ = OL.learned_match(L, R, full, 0.05, *rule)
        np.testing.assert_array_equal(tab, want)
        # random candidates, every query flagged: the fallback makes it exact
        cand = np.stack([rng.choice(512, 16, replace=False) for _ in range(640)])
        tab, fl, fr = OL.learned_match(L, R, cand, np.inf, *rule)
        assert len(fl) == 640
        np.testing.assert_array_equal(tab, want)


def test_learned_oracle_candidates_matter():
    """With random candidates and no fallback the rule decides on wrong
    top-2s: the oracle does not silently compute the exact answer."""
    d = v.synth_pair(v.synth_seed(5, 10), 512, 512, 0.4, 0.05)
    rng = np.random.default_rng(1)
    cand = np.stack([rng.choice(512, 16, replace=False) for _ in range(512)])
    tab, _, _ = OL.learned_match(d["L"], d["R"], cand, -1.0, 0.8, 0.8)
    want = O.find_u8(d["L"], d["R"], 0.8, 0.8)[2]
    assert (tab >= 0).sum() < (want >= 0).sum() / 2


def test_projection_params_make_the_network_a_projection():
    """The ablation's parameters (v3dsfmtoolkit_amd.learned.projection_params):
    through the network oracle's float64 forward pass the embedding equals
    P x / |P x| with the BF16-rounded P, since GELU(a) - GELU(-a) = a; and P
    has orthonormal rows before rounding."""
    from v3dsfmtoolkit_amd.learned import projection_params
    p = projection_params(3)
    W1 = p[:256 * 128].reshape(256, 128)
    P = W1[:64].astype(np.float64)
    np.testing.assert_allclose(P @ P.T, np.eye(64), atol=1e-6)
    U = np.random.default_rng(4).integers(0, 256, size=(200, 128)).astype(np.uint8)
    e = OM.embed(p, U)
    Pr = OM.bf16_round(W1[:64]).astype(np.float64)
    z = (U.astype(np.float64) / 64.0) @ Pr.T
    want = z / np.linalg.norm(z, axis=1, keepdims=True)
    np.testing.assert_allclose(e, want, atol=1e-12)
