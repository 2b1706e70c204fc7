"""The allocation-free inference entry point of the cost-model network.

``tensorforge_mlp_predict`` keeps the hidden activations in a file-scope device array instead of a
caller workspace. Called through the C ABI it must give bit for bit what
``tensorforge_mlp_forward`` gives with a caller workspace (same kernels, same order of
operations), on two calls in a row, the second with a smaller batch (the array keeps the first
call's contents), and refuse a batch above its limit without launching.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def _lib():
  from tensorforge_amd.runtime.loader import runtime
  rt = runtime()
  rt.tensorforge_mlp_param_count.restype = ctypes.c_long
  rt.tensorforge_mlp_param_count.argtypes = [ctypes.c_int, ctypes.c_int]
  rt.tensorforge_mlp_forward.restype = ctypes.c_int
  rt.tensorforge_mlp_forward.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
  rt.tensorforge_mlp_predict.restype = ctypes.c_int
  rt.tensorforge_mlp_predict.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
  rt.tensorforge_mlp_predict_max_activations.restype = ctypes.c_long
  rt.tensorforge_mlp_predict_max_activations.argtypes = []
  rt.tensorforge_last_error.restype = ctypes.c_int
  return rt


def _forward_with_workspace(rt, params, X, B, n_in, hidden, stream):
  import torch
  work = torch.empty(3 * B * hidden, dtype=torch.float32, device=X.device)
  out = torch.empty(B, dtype=torch.float32, device=X.device)
  assert rt.tensorforge_mlp_forward(params.data_ptr(), X.data_ptr(), work.data_ptr(),
                                    out.data_ptr(), B, n_in, hidden, stream) == 0
  return out


def test_predict_without_allocation_matches_forward(device):
  import torch
  rt = _lib()
  rng = np.random.default_rng(21)
  n_in, hidden = 27, 128
  n = rt.tensorforge_mlp_param_count(n_in, hidden)
  params = torch.from_numpy(rng.normal(0, 0.2, size=n).astype(np.float32)).to(device)
  stream = torch.cuda.current_stream().cuda_stream
  for B in (1500, 333):          # second call smaller: stale activations must not leak in
    X = torch.from_numpy(rng.normal(size=(B, n_in)).astype(np.float32)).to(device)
    ref = _forward_with_workspace(rt, params, X, B, n_in, hidden, stream)
    out = torch.full((B,), float('nan'), dtype=torch.float32, device=device)
    assert rt.tensorforge_mlp_predict(params.data_ptr(), X.data_ptr(), out.data_ptr(), B, n_in,
                                      hidden, stream) == 0
    torch.cuda.synchronize()
    got, want = out.cpu().numpy(), ref.cpu().numpy()
    print(f'B={B}: max |predict - forward| {np.abs(got - want).max():.3e}')
    assert np.array_equal(got, want)


def test_predict_refuses_batches_above_its_limit(device):
  import torch
  rt = _lib()
  hidden, n_in = 128, 4
  limit = rt.tensorforge_mlp_predict_max_activations()
  B = limit // hidden + 1
  params = torch.zeros(rt.tensorforge_mlp_param_count(n_in, hidden), dtype=torch.float32,
                       device=device)
  X = torch.zeros(B * n_in, dtype=torch.float32, device=device)
  out = torch.zeros(B, dtype=torch.float32, device=device)
  code = rt.tensorforge_mlp_predict(params.data_ptr(), X.data_ptr(), out.data_ptr(), B, n_in,
                                    hidden, torch.cuda.current_stream().cuda_stream)
  assert code != 0
  assert rt.tensorforge_last_error() == code
  rt.tensorforge_clear_error()
