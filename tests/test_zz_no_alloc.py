"""lbl_optical_depth_noalloc (no workspace argument: the scratch is a fixed
device array inside the library) against lbl_optical_depth with a caller
workspace, through the C ABI, on two calls in a row with the second one
smaller (the array keeps the first call's contents)."""

import ctypes

import numpy as np
import pytest

from pylblrtm_amd import _capi, synth

pytestmark = pytest.mark.gpu

# (lines, levels, nu_min, nu_max, step): the second case is smaller in every
# dimension, so stale records, ranges and partial sums of the first call lie
# where the second one does not write
CASES = [(3000, 13, 700.0, 712.0, 0.001), (400, 5, 720.0, 721.5, 0.0005)]


def _noalloc(prob, precision, mus, out):
    mu_arr = (ctypes.c_double * len(mus))(*[float(m) for m in mus])
    ptr = lambda k: out[k].data_ptr() if k in out else None
    c_out = _capi.LblOutputs(len(mus), mu_arr, ptr("transmittance"), ptr("optical_depth"),
                             ptr("optical_depth_layers"))
    el, eg = ctypes.c_int32(-1), ctypes.c_int32(-1)
    import torch
    stream = torch.cuda.current_stream(prob.device).cuda_stream
    with torch.cuda.device(prob.device):
        st = prob.lib.lbl_optical_depth_noalloc(
            ctypes.byref(prob.c_lines), ctypes.byref(prob.c_groups), ctypes.byref(prob.c_layers),
            ctypes.byref(prob.c_grid), _capi.PRECISIONS[precision], ctypes.byref(c_out),
            ctypes.byref(el), ctypes.byref(eg), ctypes.c_void_p(stream))
    _capi.raise_for(st, "lbl_optical_depth_noalloc")
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_workspace_entry_point(precision):
    torch = pytest.importorskip("torch")
    from pylblrtm_amd.device import OpticalDepthProblem
    dev = torch.device("cuda:0")
    mus = (1.0, 0.5)
    for n_lines, n_levels, lo, hi, step in CASES:
        print(f"noalloc {precision}: {n_lines} lines, {n_levels} levels, {lo}-{hi} @ {step}",
              flush=True)
        sd, prof = synth.co2_window(n_lines=n_lines, n_levels=n_levels)
        prob = OpticalDepthProblem(sd, prof.build_layers(), lo, hi, step, 25.0, 2.0, device=dev)
        ref = prob.run(precision, mus=mus, want_tau=True, want_layers=True)
        got = _noalloc(prob, precision, mus,
                       prob.alloc_outputs(precision, len(mus), True, True))
        torch.cuda.synchronize()
        a = ref["optical_depth_layers"].double().cpu().numpy()
        b = got["optical_depth_layers"].double().cpu().numpy()
        assert a.shape == b.shape == (prob.n_layers, prob.n_nu)
        assert np.all(a > 0)
        rel = float(np.max(np.abs(b - a) / a))
        tau_a = ref["optical_depth"].double().cpu().numpy()
        tau_b = got["optical_depth"].double().cpu().numpy()
        rel_tau = float(np.max(np.abs(tau_b - tau_a) / tau_a))
        dT = float((got["transmittance"].double() - ref["transmittance"].double()).abs().max())
        print(f"  per-layer od rel {rel:.2e}, column tau rel {rel_tau:.2e}, max |dT| {dT:.2e}",
              flush=True)
        # the same kernels on the same inputs: the two calls differ only in
        # the order of the near-centre queue's atomic additions
        if precision == "fp64":
            assert rel <= 1e-11 and rel_tau <= 1e-11 and dT <= 1e-12
        else:
            assert rel <= 1e-5 and rel_tau <= 1e-5 and dT <= 1e-6


def test_noalloc_rejects_problems_beyond_its_scratch():
    torch = pytest.importorskip("torch")
    from pylblrtm_amd.device import OpticalDepthProblem
    dev = torch.device("cuda:0")
    sd, prof = synth.co2_window(n_lines=2000, n_levels=41)
    prob = OpticalDepthProblem(sd, prof.build_layers(), 600.0, 850.0, 0.0005, 25.0, 2.0, device=dev)
    out = prob.alloc_outputs("fp64", 1, False, False)
    with pytest.raises(ValueError, match="scratch"):
        _noalloc(prob, "fp64", (1.0,), out)
