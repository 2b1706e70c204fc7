"""The autotuner's cost-model network: the FP64 NumPy oracle pinned on the CPU, the HIP
forward / backward / Adam kernels checked against it on the GPU."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _net(n_in=5, hidden=6, seed=1):
  rng = np.random.default_rng(seed)
  _, n = M.layout(n_in, hidden)
  return rng.normal(0, 0.5, size=n), rng.normal(size=(7, n_in)), rng.normal(size=7)


def test_oracle_gradient_matches_finite_differences():
  p, X, y = _net()
  g = M.backward(p, X, y, 6)
  h = 1e-6
  num = np.zeros_like(p)
  for i in range(p.size):
    d = np.zeros_like(p)
    d[i] = h
    num[i] = (M.loss(p + d, X, y, 6) - M.loss(p - d, X, y, 6)) / (2 * h)
  assert np.max(np.abs(g - num)) <= 1e-7 * max(1.0, np.max(np.abs(num)))


def test_oracle_forward_hand_computed():
  # one input, one unit per layer: out = w4 relu(w3 relu(w2 relu(w1 x + b1) + b2) + b3) + b4
  p = np.array([2.0, -1.0, 3.0, 0.5, -1.0, 4.0, 0.25, 1.0])
  out, _ = M.forward(p, np.array([[1.5]]), 1)
  a1 = max(2.0 * 1.5 - 1.0, 0)
  a2 = max(3.0 * a1 + 0.5, 0)
  a3 = max(-1.0 * a2 + 4.0, 0)
  assert out[0] == pytest.approx(0.25 * a3 + 1.0)


def test_oracle_adam_first_step_is_sign_step():
  p = np.array([1.0, -2.0, 3.0])
  g = np.array([0.5, -4.0, 1e-3])
  new, _, _ = M.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.1, eps=1e-12)
  assert np.allclose(new, p - 0.1 * np.sign(g))


@pytest.mark.gpu
def test_hip_network_matches_oracle(device):
  import torch
  from tensorforge_amd.tune.mlp import CostMLP
  rng = np.random.default_rng(3)
  n_in, B = 12, 300
  net = CostMLP(n_in, seed=4)
  X = rng.normal(size=(B, n_in)).astype(np.float32)
  y = rng.normal(size=B).astype(np.float32)
  p = net.params.double().cpu().numpy()
  Xd = torch.from_numpy(X).to(device)
  yd = torch.from_numpy(y).to(device)
  out, _ = net.forward_device(Xd)
  ref, _ = M.forward(p, X, 128)
  err = np.abs(out.double().cpu().numpy() - ref).max() / np.abs(ref).max()
  print(f'forward max rel err {err:.2e}')
  assert err <= 1e-5
  net.backward_device(Xd, yd)
  g = net.grads.double().cpu().numpy()
  gref = M.backward(p, X, y, 128)
  gerr = np.abs(g - gref).max() / np.abs(gref).max()
  print(f'gradient max rel err {gerr:.2e}')
  assert gerr <= 1e-4
  net.adam_device(1e-2)
  pnew, _, _ = M.adam(p, g, np.zeros_like(p), np.zeros_like(p), 1, lr=1e-2)
  assert np.abs(net.params.double().cpu().numpy() - pnew).max() <= 1e-6


@pytest.mark.gpu
def test_hip_training_fits_a_smooth_target(device):
  from tensorforge_amd.tune.mlp import CostMLP
  rng = np.random.default_rng(5)
  X = rng.uniform(-1, 1, size=(1024, 6))
  y = np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3] ** 2
  net = CostMLP(6, seed=6)
  hist = net.fit(X, y, epochs=60, lr=3e-3, batch=128)
  print('training loss', hist)
  assert hist[-1] < 0.05 * hist[0]
  pred = net.predict(X)
  assert np.sqrt(np.mean((pred - y) ** 2)) < 0.25 * y.std()
