"""Parity of the plain emitter's kernels with the FP64 oracle, through the C ABI on the GPU.

Tolerance (normwise per element, see ``harness.normwise_error``): 1e-12 for FP64 chains,
1e-5 for FP32 chains.
"""
import numpy as np
import pytest

from tensorforge_amd import catalog
from tensorforge_amd.common.basic_types import Addressing

from harness import TOL, gpu_run, key_of, make_inputs, normwise_error, oracle_outputs, outputs_of

pytestmark = pytest.mark.gpu

PLAIN_TESTS = [n for n in catalog.names() if n.startswith('t_') and n.endswith('_plain')]
PLAIN_WORKLOADS = ['gemm_f64_plain', 'gemm_f32_plain', 'ader_f64_plain', 'ader_f32_plain',
                   'ader_dense_f64_plain']


def _check(entry, kernel, device, num_elements, seed=0, offsets=None, flags=None):
  host = make_inputs(entry, num_elements, seed=seed, offsets=offsets)
  ref = oracle_outputs(entry, host, num_elements, offsets=offsets, flags=flags)
  got = gpu_run(kernel, entry, host, num_elements, device, offsets=offsets, flags=flags)
  tol = TOL[entry.fp_type]
  for t in entry.generator._global_tensors:
    key = key_of(t)
    if key not in outputs_of(entry) or key not in got:
      continue
    off = (offsets or {}).get(key, 0)
    if isinstance(ref[key], list):
      err = normwise_error([g[off:] for g in got[key]], [r[off:] for r in ref[key]],
                           t.get_real_volume())
    else:
      err = normwise_error(got[key], ref[key], t.get_real_volume(), off)
      # entries outside the written elements are untouched
      np.testing.assert_array_equal(got[key][:off], host[key][:off])
    print(f'{entry.name} {key} N={num_elements} normwise rel err {err:.3e} (tol {tol})')
    assert err <= tol, f'{entry.name}: {key} normwise error {err:.3e} > {tol}'


@pytest.mark.parametrize('name', PLAIN_TESTS)
def test_test_chain_matches_oracle(name, device):
  entry, kernel = catalog.load(name)
  _check(entry, kernel, device, 37)


@pytest.mark.parametrize('name', PLAIN_WORKLOADS)
def test_workload_matches_oracle(name, device):
  entry, kernel = catalog.load(name)
  _check(entry, kernel, device, 1003, seed=3)


def test_flags_and_offsets(device):
  entry, kernel = catalog.load('gemm_f64_plain')
  n = 45
  flags = (np.arange(n) % 3 != 1).astype(np.uint32)
  _check(entry, kernel, device, n, seed=5, offsets={'B': 7, 'C': 13}, flags=flags)


def test_zero_elements_is_a_no_op(device):
  entry, kernel = catalog.load('ader_f64_plain')
  host = make_inputs(entry, 4, seed=1)
  got = gpu_run(kernel, entry, host, 0, device)
  np.testing.assert_array_equal(got['Qp'], host['Qp'])
