"""Precision forms of the order-7 volume chain (84 basis functions, dense K_d) against the FP64
oracle, with degrees of freedom spanning nine decades (|Q| log-uniform in [1e-6, 1e3]).

Normwise error per element, ``max_e ||x_e - ref_e||_inf / ||ref_e||_inf``:
FP64 vector (streaming schedule, and the plain kernel) <= 1e-12; FP32 MFMA with FP32
accumulation <= 1e-5; split BF16x3 MFMA (three BF16 terms per operand, six products, a
compensating accumulator) <= 1e-5. Measured values are printed.
"""
import numpy as np
import pytest

from harness import make_inputs, normwise_error, oracle_outputs

pytestmark = pytest.mark.gpu

FORMS = {'ader7_f64_tuned': 1e-12, 'ader7_f64_plain': 1e-12, 'ader7_f64_mfma_fp32': 1e-5,
         'ader7_f64_mfma_bf16x3': 1e-5}


def _inputs(entry, n, seed=4):
  from tensorforge_amd import workloads as W
  host = make_inputs(entry, n, seed=seed)
  rng = np.random.default_rng(seed + 100)
  host['Q'] = W.log_uniform_q(rng, host['Q'].size)
  for d in range(3):
    host[f'S{d + 1}'] = host[f'S{d + 1}'] * 2.0
  return host


@pytest.mark.parametrize('name', list(FORMS))
def test_precision_form_error(name, device):
  from harness import gpu_run
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  n = 300
  host = _inputs(entry, n)
  ref = oracle_outputs(entry, host, n)['Qp']
  got = gpu_run(kernel, entry, host, n, device)['Qp']
  err = normwise_error(got, ref, entry.tensors['Qp'].get_real_volume())
  print(f'{name}: normwise rel err {err:.3e}')
  assert err <= FORMS[name]


@pytest.mark.parametrize('name', ['ader7_f64_mfma_fp32', 'ader7_f64_mfma_bf16x3'])
def test_mfma_forms_flags_and_offsets(name, device):
  from harness import gpu_run
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  n = 37
  offsets = {'Q': 3, 'Qp': 5, 'S1': 1, 'S2': 2, 'S3': 7}
  host = make_inputs(entry, n, seed=8, offsets=offsets)
  flags = (np.arange(n) % 3 != 1).astype(np.uint32)
  ref = oracle_outputs(entry, host, n, offsets=offsets, flags=flags)['Qp']
  got = gpu_run(kernel, entry, host, n, device, offsets=offsets, flags=flags)['Qp']
  vol = entry.tensors['Qp'].get_real_volume()
  assert np.array_equal(got[:5], host['Qp'][:5])          # before the offset: untouched
  for e in range(n):
    sl = slice(5 + e * vol, 5 + (e + 1) * vol)
    if flags[e]:
      assert normwise_error(got[sl], ref[sl], vol) <= 1e-5
    else:
      assert np.array_equal(got[sl], host['Qp'][sl])
