"""Batches whose per-element arrays hold more than 2^31 entries (order 6, 504 entries per element:
4,400,000 elements = 2.2e9 entries, 17.7 GB per FP64 array; about 44 GB of device memory in all).

Every element offset past 2^31 entries needs 64-bit index arithmetic in the kernels and
launchers; a 32-bit product would wrap and write the wrong element or fault, so the last elements
are the ones checked. The first two and the last four elements of the batch are copied to the host
before and after the launch and compared with the oracle (normwise per element, 1e-12 FP64); the
pointwise update with its closed form at the same points.
"""
import numpy as np
import pytest

from tensorforge_amd import catalog
from tensorforge_amd.common.basic_types import Addressing

from harness import key_of, normwise_error, oracle_outputs, outputs_of

pytestmark = pytest.mark.gpu

N = 4_400_000
SAMPLE = [0, 1, N - 4, N - 3, N - 2, N - 1]


def _need_memory(device, gib):
  import torch
  free, _ = torch.cuda.mem_get_info(device)
  if free < gib * 2 ** 30:
    pytest.skip(f'{free / 2 ** 30:.0f} GiB free, {gib} GiB needed')


def _sample(entry, bufs):
  """Host copies of the sampled elements' slices (per-element arrays) and of shared arrays."""
  host = {}
  for t in entry.generator._global_tensors:
    key = key_of(t)
    v = bufs.get(key)
    if v is None or isinstance(v, (float, int)):
      if isinstance(v, (float, int)):
        host[key] = v
      continue
    vol = t.get_real_volume()
    if t.addressing == Addressing.STRIDED:
      host[key] = np.concatenate([v[e * vol:(e + 1) * vol].cpu().numpy() for e in SAMPLE])
    elif t.addressing == Addressing.NONE:
      host[key] = v.cpu().numpy()
  return host


@pytest.mark.parametrize('name', ['ader_f64_tuned', 'ader_f64_plain'])
def test_volume_chain_past_2_31_entries(name, device):
  import torch
  from tensorforge_amd.device import launch
  from tensorforge_amd.tune.measure import synthetic_buffers
  _need_memory(device, 60)
  entry, kernel = catalog.load(name)
  assert N * 504 > 2 ** 31
  print(f'{name} N={N} ({N * 504:.3e} entries per array)', flush=True)
  bufs = synthetic_buffers(entry.generator, N, device, seed=5)
  keep = bufs.pop('__keep__')
  before = _sample(entry, bufs)
  launch(kernel, bufs, N)
  torch.cuda.synchronize()
  after = _sample(entry, bufs)
  ref = oracle_outputs(entry, before, len(SAMPLE))
  for key in outputs_of(entry):
    t = next(t for t in entry.generator._global_tensors if key_of(t) == key)
    err = normwise_error(after[key], ref[key], t.get_real_volume())
    print(f'{name} {key} at elements {SAMPLE}: normwise rel err {err:.3e}', flush=True)
    assert err <= 1e-12
  del bufs, keep
  torch.cuda.empty_cache()


def test_pointwise_update_past_2_31_entries(device):
  import torch
  from tensorforge_amd.device import launch
  _need_memory(device, 40)
  entry, kernel = catalog.load('pw_update_f64')
  g = torch.Generator(device=device).manual_seed(6)
  Q = torch.rand(N * 504, dtype=torch.float64, device=device, generator=g) - 0.5
  I = torch.rand(N * 504, dtype=torch.float64, device=device, generator=g) - 0.5
  idx = torch.cat([torch.arange(e * 504, (e + 1) * 504, device=device) for e in SAMPLE])
  q0, i0 = Q[idx].cpu().numpy(), I[idx].cpu().numpy()
  print(f'pointwise update N={N}', flush=True)
  launch(kernel, {'Q': Q, 'I': I, 'dt': 0.37}, N)
  torch.cuda.synchronize()
  got = Q[idx].cpu().numpy()
  want = q0 + 0.37 * i0
  rel = np.abs(got - want) / np.maximum(np.abs(q0) + np.abs(0.37 * i0), 1e-300)
  print(f'pointwise update at elements {SAMPLE}: max rel {rel.max():.3e}', flush=True)
  assert rel.max() <= 2.3e-16
  del Q, I
  torch.cuda.empty_cache()
