"""Learned surrogate of the Voigt function K(x, y) = Re w(x + iy).

MLP 2 -> 64 -> 64 -> 64 -> 64 -> 1 (SiLU) on u = 2 asinh(|x|)/asinh(50) - 1,
v = (log10 y + 1)/3,
trained on the GPU by the library's own BF16-MFMA forward/backward kernels
and Adam (``lbl_mlp_train_step``) on seeded samples labelled by the FP64
device Faddeeva function.  The network predicts t = log(K / K_ref) with
K_ref = exp(-x^2)/(1 + sqrt(pi) y) + y/(sqrt(pi)(x^2 + y^2 + 1/2)).

There is no counterpart in the reference (its Voigt is scipy.special.wofz,
pylblrtm/forward.py:62); the exact device Faddeeva function stays the
fallback and the yardstick of the surrogate's error.
"""

from __future__ import annotations

import ctypes
import json
import math
import time
from pathlib import Path
from typing import Optional

import numpy as np

from . import _capi

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "data" / "voigt_surrogate.npz"

# Gate of the fused surrogate: a near-centre point uses the network iff
# A * nu_c * K_ref(x, y) <= SURROGATE_GATE, i.e. its optical-depth error is at
# most TAU_BUDGET when the network's relative error is <= REL_ERR_BOUND and
# K <= 1.6 K_ref (exp(0.45), the largest log-ratio over the domain).
TAU_BUDGET = 1e-9
REL_ERR_BOUND = 0.05
SURROGATE_GATE = TAU_BUDGET / (1.6 * REL_ERR_BOUND)

_default = {}


def default_surrogate(device=None):
    """The committed network, loaded once per device."""
    key = str(device)
    if key not in _default:
        _default[key] = VoigtSurrogate.load(DEFAULT_WEIGHTS, device=device)
    return _default[key]
X_MAX = 50.0
LY_MIN, LY_MAX = -4.0, 2.0


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the Voigt surrogate runs on a ROCm GPU only")
    return torch


def k_ref(x, y):
    """Reference shape the network corrects (numpy, float64)."""
    sp = math.sqrt(math.pi)
    return np.exp(-x * x) / (1.0 + sp * y) + y / (sp * (x * x + y * y + 0.5))


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation of the parameter vector (layout of surrogate.h)."""
    rng = np.random.default_rng(seed)
    n = _capi.load().lbl_mlp_num_params()
    p = np.zeros(n, np.float32)
    H = 64
    o = 0
    p[o:o + 2 * H] = rng.normal(0, math.sqrt(2.0 / 2), 2 * H); o += 2 * H
    o += H
    for _ in range(3):
        p[o:o + H * H] = rng.normal(0, math.sqrt(2.0 / H), H * H); o += H * H
        o += H
    p[o:o + H] = rng.normal(0, math.sqrt(1.0 / H), H); o += H
    o += 1
    assert o == n
    return p


class VoigtSurrogate:
    """Device-resident parameters plus training and evaluation."""

    def __init__(self, params: Optional[np.ndarray] = None, device=None):
        torch = _torch()
        self.lib = _capi.load()
        self.device = torch.device(device if device is not None else f"cuda:{torch.cuda.current_device()}")
        if params is None:
            params = init_params()
        self.params = torch.as_tensor(np.asarray(params, np.float32)).to(self.device)
        self.wbf = torch.empty(3 * 64 * 64, dtype=torch.int16, device=self.device)
        self.meta: dict = {}
        self._pack()

    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    def _pack(self):
        st = self.lib.lbl_mlp_pack(self.params.data_ptr(), self.wbf.data_ptr(), self._stream())
        _capi.raise_for(st, "lbl_mlp_pack")

    # ---------------------------------------------------------------- data --
    def make_data(self, n: int, seed: int):
        torch = _torch()
        x = torch.empty(n, dtype=torch.float32, device=self.device)
        ly = torch.empty_like(x)
        t = torch.empty_like(x)
        st = self.lib.lbl_mlp_make_data(n, seed, x.data_ptr(), ly.data_ptr(), t.data_ptr(),
                                        self._stream())
        _capi.raise_for(st, "lbl_mlp_make_data")
        return x, ly, t

    # --------------------------------------------------------------- train --
    def gradient(self, x, ly, t, batch: int, perm=(1, 0), base: int = 0):
        """(loss sum, gradient) of one batch without updating (for checks)."""
        torch = _torch()
        grad = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        ws = self._workspace(batch)
        st = self.lib.lbl_mlp_train_step(
            self.params.data_ptr(), None, None, self.wbf.data_ptr(), x.data_ptr(), ly.data_ptr(),
            t.data_ptr(), x.numel(), perm[0], perm[1], base, batch, 0.0, 0, grad.data_ptr(),
            loss.data_ptr(), ws.data_ptr(), ws.numel(), self._stream())
        _capi.raise_for(st, "lbl_mlp_train_step")
        return float(loss.item()), grad

    def _workspace(self, batch):
        torch = _torch()
        nbytes = ctypes.c_size_t(0)
        _capi.raise_for(self.lib.lbl_mlp_train_workspace_bytes(batch, ctypes.byref(nbytes)),
                        "lbl_mlp_train_workspace_bytes")
        return torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)

    def train(self, n_samples: int = 1 << 24, batch: int = 65536, epochs: int = 20,
              lr: float = 1e-3, seed: int = 20260103, log_every: int = 0, data=None):
        """Adam on seeded, FP64-labelled samples; returns the per-epoch mean loss."""
        torch = _torch()
        if n_samples & (n_samples - 1):
            raise ValueError("n_samples must be a power of two")
        x, ly, t = data if data is not None else self.make_data(n_samples, seed)
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        grad = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        ws = self._workspace(batch)
        steps_per_epoch = n_samples // batch
        rng = np.random.default_rng(seed + 1)
        history = []
        step = 0
        t0 = time.perf_counter()
        for ep in range(epochs):
            a = int(rng.integers(0, 1 << 62)) | 1     # odd multiplier: a bijection mod 2^k
            b = int(rng.integers(0, 1 << 62))
            tot = 0.0
            for s in range(steps_per_epoch):
                step += 1
                lr_now = lr * 0.5 * (1.0 + math.cos(math.pi * (step - 1) / (epochs * steps_per_epoch)))
                st = self.lib.lbl_mlp_train_step(
                    self.params.data_ptr(), m.data_ptr(), v.data_ptr(), self.wbf.data_ptr(),
                    x.data_ptr(), ly.data_ptr(), t.data_ptr(), n_samples, a, b, s * batch, batch,
                    lr_now, step, grad.data_ptr(), loss.data_ptr(), ws.data_ptr(), ws.numel(),
                    self._stream())
                _capi.raise_for(st, "lbl_mlp_train_step")
                if s == steps_per_epoch - 1:
                    tot = float(loss.item()) / batch
            history.append(tot)
            if log_every and (ep + 1) % log_every == 0:
                print(f"[surrogate] epoch {ep + 1}/{epochs} last-batch mse {tot:.3e} "
                      f"({time.perf_counter() - t0:.1f}s)", flush=True)
        torch.cuda.synchronize(self.device)
        self.meta = dict(n_samples=n_samples, batch=batch, epochs=epochs, lr=lr, seed=seed,
                         schedule="cosine", optimizer="Adam(0.9, 0.999, 1e-8)",
                         seconds=time.perf_counter() - t0, final_mse=history[-1] if history else None)
        return history

    # ---------------------------------------------------------------- eval --
    def eval_t(self, x, ly):
        """Network output t for device tensors x, log10 y."""
        torch = _torch()
        out = torch.empty_like(x)
        st = self.lib.lbl_mlp_eval(x.numel(), x.data_ptr(), ly.data_ptr(), self.params.data_ptr(),
                                   self.wbf.data_ptr(), out.data_ptr(), self._stream())
        _capi.raise_for(st, "lbl_mlp_eval")
        return out

    def K(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """Surrogate K(x, y) = K_ref * exp(t) (numpy in/out)."""
        torch = _torch()
        xt = torch.as_tensor(np.asarray(x, np.float32)).to(self.device)
        lyt = torch.as_tensor(np.log10(np.asarray(y, np.float64)).astype(np.float32)).to(self.device)
        t = self.eval_t(xt, lyt).double().cpu().numpy()
        return k_ref(np.asarray(x, np.float64), np.asarray(y, np.float64)) * np.exp(t)

    # ------------------------------------------------------------ storage --
    def save(self, path=DEFAULT_WEIGHTS):
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, params=self.params.cpu().numpy(), meta=np.array(json.dumps(self.meta)))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device=None) -> "VoigtSurrogate":
        with np.load(path, allow_pickle=False) as f:
            s = cls(f["params"], device=device)
            s.meta = json.loads(str(f["meta"])) if "meta" in f.files else {}
        return s


def main(argv=None) -> int:
    """python -m pylblrtm_amd.surrogate [--out PATH] [--epochs N] [--samples-log2 K]"""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--samples-log2", type=int, default=24)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20260103)
    a = ap.parse_args(argv)
    s = VoigtSurrogate()
    hist = s.train(n_samples=1 << a.samples_log2, batch=a.batch, epochs=a.epochs, lr=a.lr,
                   seed=a.seed, log_every=1)
    s.meta["history"] = hist
    s.save(a.out)
    print(json.dumps({k: v for k, v in s.meta.items() if k != "history"}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
