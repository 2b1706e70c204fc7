"""The autotuner's cost model: a 3 x 128 ReLU network trained on measured kernel times.

Forward, backward and Adam run in HIP kernels of ``libtensorforge_rt.so``
(``tensorforge_mlp_*``, ``csrc/tensorforge_mlp.hip``); torch only holds the device buffers.
Inputs are standardised feature vectors; the target is log(time).
"""
import ctypes

import numpy as np

from ..runtime.loader import runtime

HIDDEN = 128


def _bind(rt):
  if getattr(rt, '_mlp_bound', False):
    return rt
  rt.tensorforge_mlp_param_count.restype = ctypes.c_long
  rt.tensorforge_mlp_param_count.argtypes = [ctypes.c_int, ctypes.c_int]
  P = ctypes.c_void_p
  rt.tensorforge_mlp_forward.restype = ctypes.c_int
  rt.tensorforge_mlp_forward.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
  rt.tensorforge_mlp_backward.restype = ctypes.c_int
  rt.tensorforge_mlp_backward.argtypes = [P, P, P, P, P, P, P, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, P]
  rt.tensorforge_mlp_adam.restype = ctypes.c_int
  rt.tensorforge_mlp_adam.argtypes = [P, P, P, P, ctypes.c_long, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_int, P]
  rt._mlp_bound = True
  return rt


def _check(code, what):
  if code != 0:
    raise RuntimeError(f'{what}: HIP error {code} ({runtime().tensorforge_error_string(code).decode()})')


def init_params(n_in, hidden=HIDDEN, seed=0):
  """He-uniform weights, zero biases, in the kernels' flat layout (host, float32)."""
  rng = np.random.default_rng(seed)
  parts = []
  for fan_in, fan_out in ((n_in, hidden), (hidden, hidden), (hidden, hidden), (hidden, 1)):
    lim = np.sqrt(6.0 / fan_in)
    parts.append(rng.uniform(-lim, lim, size=fan_out * fan_in))
    parts.append(np.zeros(fan_out))
  return np.concatenate(parts).astype(np.float32)


class CostMLP:
  """``fit(X, y)`` / ``predict(X)`` on host arrays; the arithmetic runs on ``device``."""

  def __init__(self, n_in, hidden=HIDDEN, seed=0, device='cuda:0'):
    import torch
    self.torch = torch
    self.rt = _bind(runtime())
    self.n_in, self.hidden = n_in, hidden
    self.device = torch.device(device)
    n = self.rt.tensorforge_mlp_param_count(n_in, hidden)
    host = init_params(n_in, hidden, seed)
    assert host.size == n
    self.params = torch.from_numpy(host).to(self.device)
    self.m = torch.zeros_like(self.params)
    self.v = torch.zeros_like(self.params)
    self.grads = torch.zeros_like(self.params)
    self.step = 0
    self.mean = np.zeros(n_in, dtype=np.float32)
    self.std = np.ones(n_in, dtype=np.float32)
    self.y_mean, self.y_std = 0.0, 1.0
    self._ws = {}

  # ---- raw kernels on device tensors ----
  def _stream(self):
    return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

  def _buffers(self, B):
    """Activation, gradient and output buffers for batches of B rows (kept per B: training
    reuses the same few sizes every epoch)."""
    bufs = self._ws.get(B)
    if bufs is None:
      t = self.torch
      bufs = (t.empty(3 * B * self.hidden, dtype=t.float32, device=self.device),
              t.empty(2 * B * self.hidden, dtype=t.float32, device=self.device),
              t.empty(B, dtype=t.float32, device=self.device))
      self._ws[B] = bufs
    return bufs

  def forward_device(self, X, fresh=True):
    B = X.shape[0]
    if fresh:
      work = self.torch.empty(3 * B * self.hidden, dtype=self.torch.float32, device=self.device)
      out = self.torch.empty(B, dtype=self.torch.float32, device=self.device)
    else:
      work, _, out = self._buffers(B)
    _check(self.rt.tensorforge_mlp_forward(self.params.data_ptr(), X.data_ptr(), work.data_ptr(),
                                           out.data_ptr(), B, self.n_in, self.hidden,
                                           self._stream()), 'tensorforge_mlp_forward')
    return out, work

  def backward_device(self, X, target):
    out, work = self.forward_device(X, fresh=False)
    B = X.shape[0]
    dwork = self._buffers(B)[1]
    _check(self.rt.tensorforge_mlp_backward(self.params.data_ptr(), X.data_ptr(), work.data_ptr(),
                                            out.data_ptr(), target.data_ptr(),
                                            self.grads.data_ptr(), dwork.data_ptr(), B, self.n_in,
                                            self.hidden, self._stream()),
           'tensorforge_mlp_backward')
    return out

  def adam_device(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    self.step += 1
    _check(self.rt.tensorforge_mlp_adam(self.params.data_ptr(), self.grads.data_ptr(),
                                        self.m.data_ptr(), self.v.data_ptr(), self.params.numel(),
                                        lr, beta1, beta2, eps, self.step, self._stream()),
           'tensorforge_mlp_adam')

  # ---- training on host arrays ----
  def _x(self, X):
    Xs = (np.asarray(X, dtype=np.float32) - self.mean) / self.std
    return self.torch.from_numpy(np.ascontiguousarray(Xs, dtype=np.float32)).to(self.device)

  def fit(self, X, y, epochs=300, lr=2e-3, batch=256, seed=0):
    X = np.asarray(X, dtype=np.float32)
    y = np.asarray(y, dtype=np.float64)
    self.mean = X.mean(axis=0)
    self.std = np.where(X.std(axis=0) > 1e-8, X.std(axis=0), 1.0).astype(np.float32)
    self.y_mean, self.y_std = float(y.mean()), float(y.std() or 1.0)
    Xd = self._x(X)
    yd = self.torch.from_numpy(((y - self.y_mean) / self.y_std).astype(np.float32)).to(self.device)
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    history = []
    for ep in range(epochs):
      # one shuffled copy per epoch; the minibatches are contiguous slices of it
      order = self.torch.from_numpy(rng.permutation(n)).to(self.device)
      Xe = Xd.index_select(0, order)
      ye = yd.index_select(0, order)
      for s in range(0, n, batch):
        self.backward_device(Xe[s:s + batch], ye[s:s + batch])
        self.adam_device(lr)
      if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
        out, _ = self.forward_device(Xd)
        history.append(float(((out - yd) ** 2).mean().item()))
    return history

  def predict(self, X):
    out, _ = self.forward_device(self._x(X))
    return out.double().cpu().numpy() * self.y_std + self.y_mean
