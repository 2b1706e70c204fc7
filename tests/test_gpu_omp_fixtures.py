"""This package's GPU kernels (plain and default emitter) on the inputs of the OpenMP golden
fixtures, against the outputs the reference's own omptarget-generated kernels wrote
(tests/golden/omp_fixtures.{json,npz}), through the C ABI. Normwise per element: <= 1e-12 for
FP64 kernels, <= 1e-5 for FP32 kernels (the reference computes those in FP32 too)."""
import numpy as np
import pytest

import omp_fixture_build as fb
import omp_fixture_cases as fx
from harness import normwise_error

pytestmark = pytest.mark.gpu

ZEROS = {'gemm_descr_alpha'}     # the reference kernel writes zeros there (see test_omp_fixtures)
CASES = [(r['label'], e) for r in fx.CASES if r['label'] not in ZEROS for e in fb.EMITTERS]


@pytest.mark.parametrize('label,emitter', CASES)
def test_kernel_matches_reference_omp_kernel(label, emitter, device):
  import torch
  from tensorforge_amd.device import launch
  from tensorforge_amd.runtime.loader import load_kernels
  rec = next(r for r in fx.CASES if r['label'] == label)
  gen, _, tensors = fx.generator(rec, emitter)
  kernel = load_kernels(fb.so_path(label, emitter), [gen])[0]
  n = rec['elements']
  print(f'{label} {emitter} ({gen.get_emitted().emitter}) N={n}', flush=True)
  host = fx.host_arrays(rec, tensors)
  bufs = {}
  for alias, v in host.items():
    if isinstance(v, list):
      bufs[alias] = [torch.from_numpy(x.copy()).to(device) for x in v]
    elif isinstance(v, float):
      bufs[alias] = v
    else:
      bufs[alias] = torch.from_numpy(v.copy()).to(device)
  keep = launch(kernel, bufs, n)
  torch.cuda.synchronize()
  del keep
  tol = 1e-12 if rec['fp_type'] == 'double' else 1e-5
  for alias in rec['outputs']:
    p = next(p for p in rec['params'] if p['alias'] == alias)
    got = bufs[alias]
    got = np.concatenate([x.cpu().numpy() for x in got]) if isinstance(got, list) else \
        got.cpu().numpy()
    err = normwise_error(got, fx.expected(rec, alias), p['volume'])
    print(f'  {alias}: normwise {err:.2e} (tol {tol})')
    assert err <= tol
