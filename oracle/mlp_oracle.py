"""FP64 NumPy restatement of the autotuner's cost-model network, its mean-squared-error
gradient and its Adam step.

TEST INFRASTRUCTURE ONLY: the checker the HIP network kernels (libtensorforge_rt.so,
``tensorforge_mlp_*``) are compared against; only ``tests/`` may use it.

The network has no counterpart in the reference. Its definition (fully connected, ReLU hidden
layers, parameters ``W1 b1 W2 b2 W3 b3 W4 b4`` flattened row-major in that order) and Adam
(Kingma & Ba, 2015, with bias correction) are the textbook ones. Pinned in ``tests/test_mlp.py``
by central finite differences of the loss (gradient), a hand-computed one-unit network
(forward), and the closed form of Adam's first step (``-lr * sign(g)`` when eps is negligible).
"""
import numpy as np


def layout(n_in, hidden):
  sizes = [('W1', (hidden, n_in)), ('b1', (hidden,)), ('W2', (hidden, hidden)), ('b2', (hidden,)),
           ('W3', (hidden, hidden)), ('b3', (hidden,)), ('W4', (1, hidden)), ('b4', (1,))]
  out, pos = {}, 0
  for name, shape in sizes:
    n = int(np.prod(shape))
    out[name] = (pos, shape)
    pos += n
  return out, pos


def unpack(params, n_in, hidden):
  lay, _ = layout(n_in, hidden)
  return {k: np.asarray(params[p:p + int(np.prod(s))], dtype=np.float64).reshape(s)
          for k, (p, s) in lay.items()}


def forward(params, X, hidden):
  """Returns (out[B], activations [a1, a2, a3])."""
  X = np.asarray(X, dtype=np.float64)
  P = unpack(params, X.shape[1], hidden)
  a1 = np.maximum(X @ P['W1'].T + P['b1'], 0.0)
  a2 = np.maximum(a1 @ P['W2'].T + P['b2'], 0.0)
  a3 = np.maximum(a2 @ P['W3'].T + P['b3'], 0.0)
  out = (a3 @ P['W4'].T + P['b4'])[:, 0]
  return out, [a1, a2, a3]


def loss(params, X, target, hidden):
  out, _ = forward(params, X, hidden)
  return float(np.mean((out - np.asarray(target, dtype=np.float64)) ** 2))


def backward(params, X, target, hidden):
  """Gradient of mean((out - target)^2), flattened like ``params``."""
  X = np.asarray(X, dtype=np.float64)
  B = X.shape[0]
  P = unpack(params, X.shape[1], hidden)
  out, (a1, a2, a3) = forward(params, X, hidden)
  dout = 2.0 * (out - np.asarray(target, dtype=np.float64)) / B
  g = {}
  g['W4'] = dout[None, :] @ a3
  g['b4'] = np.array([dout.sum()])
  dz3 = (dout[:, None] @ P['W4']) * (a3 > 0)
  g['W3'] = dz3.T @ a2
  g['b3'] = dz3.sum(axis=0)
  dz2 = (dz3 @ P['W3']) * (a2 > 0)
  g['W2'] = dz2.T @ a1
  g['b2'] = dz2.sum(axis=0)
  dz1 = (dz2 @ P['W2']) * (a1 > 0)
  g['W1'] = dz1.T @ X
  g['b1'] = dz1.sum(axis=0)
  lay, n = layout(X.shape[1], hidden)
  flat = np.zeros(n)
  for k, (p, s) in lay.items():
    flat[p:p + int(np.prod(s))] = g[k].reshape(-1)
  return flat


def adam(params, grads, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
  """One bias-corrected Adam step; returns (params, m, v)."""
  m = beta1 * m + (1 - beta1) * grads
  v = beta2 * v + (1 - beta2) * grads * grads
  mh = m / (1 - beta1 ** step)
  vh = v / (1 - beta2 ** step)
  return params - lr * mh / (np.sqrt(vh) + eps), m, v
