"""Voigt surrogate: MFMA forward/backward kernels against the FP64 MLP
oracle, labels against scipy, training progress, and the trained network's
accuracy.  Needs a ROCm GPU."""

import math
import sys

import numpy as np
import pytest
from scipy.special import wofz

from oracle import mlp_oracle as M
from pylblrtm_amd import surrogate as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rand_params(seed):
    p = S.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    p += rng.normal(0, 0.05, p.size)          # non-zero biases too
    return p.astype(np.float32)


def test_labels_match_scipy():
    say("case: surrogate labels")
    s = S.VoigtSurrogate(_rand_params(1))
    x, ly, t = s.make_data(1 << 14, 123)
    x, ly, t = (a.double().cpu().numpy() for a in (x, ly, t))
    assert abs(x).max() <= 50.0 and ly.min() >= -4.0 and ly.max() <= 2.0
    y = 10.0 ** ly
    ref = np.log(wofz(x + 1j * y).real / S.k_ref(x, y))
    assert np.max(np.abs(t - ref)) < 2e-6
    assert -0.75 < t.min() and t.max() < 0.5


def test_mfma_forward_matches_oracle():
    say("case: surrogate forward vs oracle")
    import torch
    p = _rand_params(2)
    s = S.VoigtSurrogate(p)
    rng = np.random.default_rng(5)
    x = rng.uniform(-50, 50, 1000).astype(np.float32)      # not a multiple of 64
    ly = rng.uniform(-4, 2, 1000).astype(np.float32)
    t = s.eval_t(torch.as_tensor(x, device=s.device), torch.as_tensor(ly, device=s.device))
    ref = M.forward(p, x, ly)[0]
    assert np.max(np.abs(t.double().cpu().numpy() - ref)) < 5e-3 * max(1.0, np.abs(ref).max())


def test_training_gradient_matches_oracle():
    say("case: surrogate gradient vs oracle")
    p = _rand_params(3)
    s = S.VoigtSurrogate(p)
    x, ly, t = s.make_data(1 << 12, 77)
    loss, g = s.gradient(x, ly, t, batch=512, perm=(1, 0), base=0)
    xs, ls, ts = (a[:512].double().cpu().numpy() for a in (x, ly, t))
    ref_loss, ref_g = M.loss_and_grad(p, xs.astype(np.float32), ls.astype(np.float32), ts)
    assert loss == pytest.approx(ref_loss, rel=2e-3)
    gg = g.double().cpu().numpy()
    assert np.linalg.norm(gg - ref_g) / np.linalg.norm(ref_g) < 2e-2


def test_training_reduces_loss():
    say("case: surrogate short training")
    s = S.VoigtSurrogate(S.init_params(4))
    data = s.make_data(1 << 18, 11)
    loss0 = s.gradient(*data, batch=4096)[0] / 4096
    hist = s.train(n_samples=1 << 18, batch=4096, epochs=6, lr=2e-3, seed=11, data=data)
    say(f"untrained mse {loss0:.3e}, per-epoch {hist}")
    assert hist[-1] < 0.05 * loss0 and hist[-1] < hist[0], (loss0, hist)


def test_committed_network_accuracy():
    say("case: committed surrogate accuracy")
    if not S.DEFAULT_WEIGHTS.exists():
        pytest.skip("no trained weights committed")
    s = S.VoigtSurrogate.load()
    rng = np.random.default_rng(2027)
    x = rng.uniform(-50, 50, 1 << 16)
    y = 10.0 ** rng.uniform(-4, 2, 1 << 16)
    K = s.K(x, y)
    ref = wofz(x + 1j * y).real
    rel = np.abs(K - ref) / ref
    say(f"surrogate rel err: median {np.median(rel):.2e}, p99 {np.quantile(rel, 0.99):.2e}, "
        f"max {rel.max():.2e}")
    assert np.quantile(rel, 0.99) < 2e-2 and rel.max() < 1e-1
