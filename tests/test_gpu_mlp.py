"""GPU training and inference kernels of the embedding network against the
float64 CPU oracle (oracle/mlp.py): batch loss, every parameter gradient, and
embeddings; plus one short training run."""
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def pair_batch(seed, B):
    d = v.synth_pair(v.synth_seed(6, seed), 2 * B, 2 * B, 1.0, 0.05)
    src = d["src"]
    j = np.flatnonzero(src >= 0)[:B]
    return d["L"][src[j]], d["R"][j]


@pytest.fixture(scope="module")
def net(gpu_available):
    n = v.EmbeddingNet(0, seed=3)
    yield n
    n.close()


@pytest.mark.parametrize("B,tau", [(64, 0.05), (200, 0.1), (512, 0.05)])
def test_loss_and_gradients_match_oracle(net, B, tau):
    say("mlp grad", B, tau)
    rng = np.random.default_rng(B)
    p = np.zeros(OM.N_PARAMS, np.float32)
    p[OM.OFF_W1:OM.OFF_B1] = rng.normal(0, np.sqrt(2 / 128), OM.OFF_B1)
    p[OM.OFF_B1:OM.OFF_W2] = rng.normal(0, 0.1, 256)
    p[OM.OFF_W2:OM.OFF_B2] = rng.normal(0, np.sqrt(1 / 256), 64 * 256)
    p[OM.OFF_B2:] = rng.normal(0, 0.1, 64)
    net.set_params(p)
    np.testing.assert_array_equal(net.params(), p)
    UL, UR = pair_batch(B, B)
    loss, g = net.loss_grad(UL, UR, tau)
    want_loss, want_g = OM.loss_grad(p, UL, UR, tau)
    assert loss == pytest.approx(want_loss, rel=2e-5, abs=2e-6)
    for name, a, b in [("W1", OM.OFF_W1, OM.OFF_B1), ("b1", OM.OFF_B1, OM.OFF_W2), ("W2", OM.OFF_W2, OM.OFF_B2),
                       ("b2", OM.OFF_B2, OM.N_PARAMS)]:
        x, y = g[a:b].astype(np.float64), want_g[a:b]
        err = np.abs(x - y).max() / max(np.abs(y).max(), 1e-12)
        say(f"  {name}: max |gpu - oracle| / max |oracle| = {err:.2e}")
        assert err < 1e-4, name


def test_embeddings_match_oracle(net):
    say("mlp embed")
    rng = np.random.default_rng(5)
    p = np.zeros(OM.N_PARAMS, np.float32)
    p[OM.OFF_W1:OM.OFF_B1] = rng.normal(0, np.sqrt(2 / 128), OM.OFF_B1)
    p[OM.OFF_W2:OM.OFF_B2] = rng.normal(0, np.sqrt(1 / 256), 64 * 256)
    net.set_params(p)
    d = v.synth_pair(v.synth_seed(6, 99), 3000, 32)
    got = net.embed(d["L"])
    want = OM.embed(p, d["L"])
    assert np.abs(got - want).max() < 2e-5
    np.testing.assert_allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-5)


def test_short_training_lowers_loss(gpu_available):
    say("mlp train")
    from v3dsfmtoolkit_amd import learned as VL
    L, R = VL.labeled_pairs(20000, n=4096)
    net = v.EmbeddingNet(0, seed=11)
    try:
        held = slice(0, 512)
        before, _ = net.loss_grad(L[held], R[held], 0.05)
        losses = VL.train(net, L[512:], R[512:], steps=150, batch=512, lr=2e-3)
        after, _ = net.loss_grad(L[held], R[held], 0.05)
        say(f"held-out loss {before:.4f} -> {after:.4f}; batch loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        assert np.all(np.isfinite(losses))
        assert after < 0.5 * before
    finally:
        net.close()
