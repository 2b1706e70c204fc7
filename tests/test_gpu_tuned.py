"""Parity of the tuned emitter's kernels with the FP64 oracle, through the C ABI on the GPU.

Tolerance (normwise per element, see ``harness.normwise_error``): 1e-12 for FP64 chains,
1e-5 for FP32 chains.
"""
import numpy as np
import pytest

from tensorforge_amd import catalog

from harness import TOL, gpu_run, key_of, make_inputs, normwise_error, oracle_outputs, outputs_of
from test_gpu_plain import _check

pytestmark = pytest.mark.gpu

TUNED_WORKLOADS = ['gemm_f64_tuned', 'gemm_f32_tuned', 'ader_f64_tuned', 'ader_f32_tuned',
                   'ader_dense_f64_tuned', 'ader_dense_f32_tuned']
AUTO_TESTS = [n for n in catalog.names() if n.startswith('t_') and n.endswith('_auto')]


@pytest.mark.parametrize('name', TUNED_WORKLOADS)
def test_tuned_workload_matches_oracle(name, device):
  entry, kernel = catalog.load(name)
  assert entry.generator.get_emitted().emitter == 'tuned'
  _check(entry, kernel, device, 1003, seed=3)


@pytest.mark.parametrize('name', AUTO_TESTS)
def test_auto_chain_matches_oracle(name, device):
  entry, kernel = catalog.load(name)
  print(name, 'emitter', entry.generator.get_emitted().emitter)
  _check(entry, kernel, device, 37)


@pytest.mark.parametrize('n', [1, 6, 7, 8, 64, 65])
def test_tuned_partial_groups(n, device):
  entry, kernel = catalog.load('ader_f64_tuned')
  _check(entry, kernel, device, n, seed=n)


def test_tuned_flags_and_offsets(device):
  entry, kernel = catalog.load('ader_f64_tuned')
  n = 50
  flags = (np.arange(n) % 4 != 2).astype(np.uint32)
  _check(entry, kernel, device, n, seed=9, offsets={'Q': 3, 'Qp': 11, 'S2': 5}, flags=flags)


def test_tuned_agrees_with_plain(device):
  """Both emitters on the same inputs: the same chain up to FP64 rounding."""
  e1, plain = catalog.load('ader_f64_plain')
  e2, tuned = catalog.load('ader_f64_tuned')
  n = 700
  host = make_inputs(e1, n, seed=21)
  a = gpu_run(plain, e1, host, n, device)['Qp']
  b = gpu_run(tuned, e2, host, n, device)['Qp']
  err = normwise_error(b, a, 504)
  print(f'tuned vs plain normwise rel diff {err:.3e}')
  assert err < 1e-13
