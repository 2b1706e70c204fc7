"""The headline chain at its full batch (order 6, 262,144 elements, FP64 and FP32): checks whose
cost does not grow with the batch on the CPU side.

* a seeded sample of 512 elements (spread over the whole batch) against the FP64 oracle;
* tuned = plain kernel (dense-K form of the same chain) on every element;
* linearity in the streamed operand: f(Q1 + 2 Q2) = f(Q1) + 2 f(Q2) on every element.
"""
import numpy as np
import pytest

from harness import TOL, normwise_error, oracle_outputs

pytestmark = pytest.mark.gpu

N = 262144


def _bufs(entry, device, dtype, seed):
  import torch
  from tensorforge_amd import workloads as W
  g = torch.Generator(device=device)
  g.manual_seed(seed)
  b, q = entry.meta['b'], entry.meta['q']
  out = {}
  for d in range(3):
    nnz = W.packed_count(entry.tensors[f'K{d + 1}'])
    out[f'K{d + 1}'] = (torch.rand(nnz, generator=g, device=device, dtype=torch.float64) * 2 - 1).to(dtype)
    s = torch.zeros(N, q * q, device=device, dtype=torch.float64)
    s[:, :24] = torch.rand(N, 24, generator=g, device=device, dtype=torch.float64) * 4 - 2
    out[f'S{d + 1}'] = s.reshape(-1).to(dtype)
  out['Q'] = torch.randn(N * b * q, generator=g, device=device, dtype=torch.float64).to(dtype)
  out['Qp'] = torch.zeros(N * b * q, device=device, dtype=dtype)
  return out


def _run(kernel, bufs):
  import torch
  from tensorforge_amd.device import launch
  launch(kernel, bufs, N)
  torch.cuda.synchronize()
  return bufs['Qp'].double().cpu().numpy()


@pytest.mark.parametrize('name', ['ader_f64_tuned', 'ader_f32_tuned'])
def test_full_batch_sample_against_oracle(name, device):
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  dtype = __import__('torch').float64 if 'f64' in name else __import__('torch').float32
  bufs = _bufs(entry, device, dtype, seed=31)
  got = _run(kernel, bufs)
  vol = 56 * 9
  idx = np.sort(np.random.default_rng(32).choice(N, 512, replace=False))
  host = {}
  for key, v in bufs.items():
    a = v.double().cpu().numpy()
    if key.startswith('K'):
      host[key] = a
    elif key.startswith('S'):
      host[key] = a.reshape(N, 81)[idx].reshape(-1)
    else:
      host[key] = a.reshape(N, vol)[idx].reshape(-1)
  host['Qp'] = np.zeros(512 * vol)
  ref = oracle_outputs(entry, host, 512)['Qp']
  err = normwise_error(got.reshape(N, vol)[idx].reshape(-1), ref, vol)
  print(f'{name}: sample of 512 elements, normwise rel err {err:.2e}')
  assert err <= TOL[entry.fp_type]


def test_full_batch_tuned_equals_plain_dense_and_linearity(device):
  import torch
  from tensorforge_amd import catalog
  from tensorforge_amd import workloads as W
  entry, tuned = catalog.load('ader_f64_tuned')
  _, plain = catalog.load('ader_dense_f64_plain')
  a = _bufs(entry, device, torch.float64, seed=41)
  got = _run(tuned, a)
  # the same arithmetic with K densified, through the plain kernel
  d = dict(a)
  for k in range(3):
    dense = torch.zeros(56 * 56, dtype=torch.float64, device=device)
    ents = W.csr_entries(entry.meta['masks'][k])
    dense[torch.tensor([r + 56 * c for r, c in ents], device=device)] = a[f'K{k + 1}']
    d[f'K{k + 1}'] = dense
  d['Qp'] = torch.zeros_like(a['Qp'])
  ref = _run(plain, d)
  err = normwise_error(got, ref, 56 * 9)
  print(f'tuned vs plain (dense K) on {N} elements: {err:.2e}')
  assert err <= 1e-13
  # linearity in Q
  b = dict(a)
  b['Q'] = torch.randn(a['Q'].shape, generator=torch.Generator(device=device).manual_seed(42),
                       device=device, dtype=torch.float64)
  b['Qp'] = torch.zeros_like(a['Qp'])
  fb = _run(tuned, b)
  c = dict(a)
  c['Q'] = a['Q'] + 2 * b['Q']
  c['Qp'] = torch.zeros_like(a['Qp'])
  fc = _run(tuned, c)
  lin = normwise_error(fc, got + 2 * fb, 56 * 9)
  print(f'linearity on {N} elements: {lin:.2e}')
  assert lin <= 1e-12
