"""Kernels generated through yateto's legacy gateway (tests/golden/yateto_cases.py: pointer-based
per-element operands, aligned and unaligned windows, the ADER chain with sparse star and K
matrices, a view operand), run through the C ABI on the GPU and compared with the FP64 oracle
interpreting the frontend's description list. Normwise per element <= 1e-12 (FP64); entries
outside the written window come back unchanged."""
import numpy as np
import pytest

import yateto_gateway_build as yb
import yateto_cases as yc
from harness import normwise_error
from oracle.chain_oracle import OracleInputs, run as oracle_run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('case', list(yc.CASES))
def test_gateway_kernel_matches_oracle(case, device):
  import torch
  from tensorforge_amd.common.basic_types import Addressing
  from tensorforge_amd.device import launch
  from tensorforge_amd.runtime.loader import load_kernels
  fe, gens = yb.generators(case)
  assert len(gens) == 1
  gen = gens[0]
  kernel = load_kernels(yb.so_path(case), [gen])[0]
  n = 77
  rng = np.random.default_rng(31)
  host, arrays = {}, {}
  for t in gen._global_tensors:
    vol = t.get_real_volume()
    if t.addressing == Addressing.NONE:
      v = rng.uniform(-1, 1, t.spp.count_nz())
      host[t.alias], arrays[t] = v, v.copy()
    elif t.addressing == Addressing.STRIDED:
      v = rng.uniform(-1, 1, n * vol)
      host[t.alias], arrays[t] = v, v.copy()
    else:
      v = [rng.uniform(-1, 1, vol) for _ in range(n)]
      host[t.alias], arrays[t] = v, [x.copy() for x in v]
  want = oracle_run(fe.generator._descr_list, OracleInputs(arrays, n))
  bufs = {k: ([torch.from_numpy(x.copy()).to(device) for x in v] if isinstance(v, list)
              else torch.from_numpy(v.copy()).to(device)) for k, v in host.items()}
  print(f'{case}: {gen.get_base_name()} ({gen.get_emitted().emitter}) N={n}', flush=True)
  keep = launch(kernel, bufs, n)
  torch.cuda.synchronize()
  del keep
  for t in gen._global_tensors:
    if t not in want or t.addressing == Addressing.NONE:
      continue
    got = bufs[t.alias]
    got = [x.cpu().numpy() for x in got] if isinstance(got, list) else got.cpu().numpy()
    ref = want[t]
    err = normwise_error(got, ref, t.get_real_volume())
    print(f'  {t.alias}: normwise {err:.2e}')
    assert err <= 1e-12
