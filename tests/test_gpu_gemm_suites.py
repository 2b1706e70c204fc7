"""The reference's GEMM test cases (tests/gemm_suites.py: 514 dense cases, the sparse-operand
suites, the scale-add suites) on the GPU through the C ABI, plain and auto emitters, against the
oracle. Normwise per element <= 1e-12 (FP64); entries of C outside the computed window must come
back unchanged."""
import os

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


def _run_against_oracle(kernel, descrs, tensors, host, n, device, run, label):
  import torch
  from tensorforge_amd.device import launch
  inp = OracleInputs({tensors[k]: host[k].copy() for k in host}, n)
  want = run(descrs, inp)[tensors['C']]
  bufs = {k: torch.from_numpy(v.copy()).to(device) for k, v in host.items()}
  launch(kernel, bufs, n)
  torch.cuda.synchronize()
  got = bufs['C'].cpu().numpy().reshape(n, -1)
  w = want.reshape(n, -1)
  err = float((np.abs(got - w).max(axis=1) / np.maximum(np.abs(w).max(axis=1), 1e-300)).max())
  print(f'{label}: normwise error {err:.3e}', flush=True)
  assert err <= 1e-12, label


@pytest.mark.parametrize('emitter', gs.EMITTERS)
def test_sparse_suites_match_oracle(emitter, device):
  from tensorforge_amd.runtime.loader import load_kernels
  from test_gemm_suites import sparse_host
  cases = gs.sparse_cases()
  built = [gs.sparse_generator(c, emitter) for c in cases]
  kernels = [load_kernels(os.path.join(gs.BUILD_DIR, f'libsparse_suites_{emitter}_{i}.so'), [g])[0]
             for i, (g, _, _) in enumerate(built)]
  for i, (case, (g, descrs, tensors), k) in enumerate(zip(cases, built, kernels)):
    print(f'{case["name"]} {emitter} ({g.get_emitted().emitter}) N={case["n"]}', flush=True)
    host, _ = sparse_host(case, i)
    _run_against_oracle(k, descrs, tensors, host, case['n'], device, oracle_run,
                        f'{case["name"]} {emitter}')


def test_csa_suites_match_oracle(device):
  from oracle.pointwise_oracle import run as pw_run
  from tensorforge_amd.runtime.loader import load_kernels
  from test_gemm_suites import csa_host
  cases = gs.csa_cases()
  built = [gs.csa_generator(c) for c in cases]
  kernels = load_kernels(os.path.join(gs.BUILD_DIR, 'libcsa_suites.so'), [g for g, _, _ in built])
  for i, (case, (g, descrs, tensors), k) in enumerate(zip(cases, built, kernels)):
    print(f'{case["name"]} alpha {case["alpha"]} beta {case["beta"]} N={case["n"]}', flush=True)
    _run_against_oracle(k, descrs, tensors, csa_host(case, i), case['n'], device, pw_run,
                        case['name'])
