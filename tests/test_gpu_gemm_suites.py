"""The reference's 514 dense GEMM test cases (tests/gemm_suites.py) on the GPU through the C ABI,
plain and auto emitters, against the oracle. Normwise per element <= 1e-12 (FP64); entries of C
outside the computed window must come back unchanged."""
import numpy as np
import pytest

import gemm_suites as gs
from oracle.chain_oracle import OracleInputs, run as oracle_run
from test_gemm_suites import host_inputs, offsets

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('emitter', gs.EMITTERS)
@pytest.mark.parametrize('part', range(gs.PARTS))
def test_suite_part_matches_oracle(emitter, part, device):
  import torch
  from tensorforge_amd.device import launch
  from tensorforge_amd.runtime.loader import load_kernels
  chosen = [(i, c) for i, c in enumerate(gs.cases()) if gs.part_of(i) == part]
  built = [gs.generator(c, emitter) for _, c in chosen]
  kernels = load_kernels(gs.so_path(emitter, part), [g for g, _, _ in built])
  worst = 0.0
  for (index, case), (gen, descrs, tensors), kernel in zip(chosen, built, kernels):
    print(f'{case["name"]} {emitter} N={case["n"]}', flush=True)
    host = host_inputs(case, index)
    offs = offsets(case)
    inp = OracleInputs({tensors[k]: host[k].copy() for k in host}, case['n'],
                       offsets={tensors[k]: offs[k] for k in offs})
    want = oracle_run(descrs, inp)[tensors['C']]
    bufs = {k: torch.from_numpy(v.copy()).to(device) for k, v in host.items()}
    launch(kernel, bufs, case['n'], offsets=offs)
    torch.cuda.synchronize()
    got = bufs['C'].cpu().numpy()
    vol = case['c'][0] * case['c'][1]
    g = got.reshape(case['n'], vol)
    w = want.reshape(case['n'], vol)
    err = float((np.abs(g - w).max(axis=1) / np.maximum(np.abs(w).max(axis=1), 1e-300)).max())
    worst = max(worst, err)
    assert err <= 1e-12, f'{case["name"]} ({emitter}): normwise error {err:.3e}'
  print(f'{len(chosen)} cases, {emitter}, worst normwise error {worst:.3e}', flush=True)
