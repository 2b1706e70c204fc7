"""The FP64 matrix-core form (``emitter='mfma_fp64'``: ``v_mfma_f64_16x16x4f64`` tiles of the dense
shared operand, FP64 throughout) against the FP64 oracle, at the FP64 bound (normwise 1e-12):
the single GEMM ``C += A B`` (56x56 shared, 56x9 per element) and the order-6 volume chain with
dense K_d (its temporaries ``Q S_d`` formed on the vector units). Ragged batches (not a multiple of
the 7 elements a wavefront owns), flags and extra offsets, and a many-round batch checked
against the streaming form of the same chain.
"""
import numpy as np
import pytest

from harness import gpu_run, make_inputs, normwise_error, oracle_outputs

pytestmark = pytest.mark.gpu

FORMS = {'gemm_f64_mfma_fp64': 'C', 'ader_dense_f64_mfma_fp64': 'Qp'}
OFFSETS = {'gemm_f64_mfma_fp64': {'B': 3, 'C': 5},
           'ader_dense_f64_mfma_fp64': {'Q': 3, 'Qp': 5, 'S1': 1, 'S2': 2, 'S3': 7}}


@pytest.mark.parametrize('name', list(FORMS))
def test_fp64_mfma_matches_oracle(name, device):
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  assert entry.generator.get_emitted().schedule['kind'] == 'mfma_fp64'
  out = FORMS[name]
  for n in (1, 300):
    host = make_inputs(entry, n, seed=5)
    ref = oracle_outputs(entry, host, n)[out]
    got = gpu_run(kernel, entry, host, n, device)[out]
    err = normwise_error(got, ref, entry.tensors[out].get_real_volume())
    print(f'{name} N={n}: normwise rel err {err:.3e}')
    assert err <= 1e-12


@pytest.mark.parametrize('name', list(FORMS))
def test_fp64_mfma_flags_and_offsets(name, device):
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  out = FORMS[name]
  n = 37
  offsets = OFFSETS[name]
  host = make_inputs(entry, n, seed=8, offsets=offsets)
  flags = (np.arange(n) % 3 != 1).astype(np.uint32)
  ref = oracle_outputs(entry, host, n, offsets=offsets, flags=flags)[out]
  got = gpu_run(kernel, entry, host, n, device, offsets=offsets, flags=flags)[out]
  vol = entry.tensors[out].get_real_volume()
  off = offsets[out]
  assert np.array_equal(got[:off], host[out][:off])           # before the offset: untouched
  for e in range(n):
    sl = slice(off + e * vol, off + (e + 1) * vol)
    if flags[e]:
      assert normwise_error(got[sl], ref[sl], vol) <= 1e-12
    else:
      assert np.array_equal(got[sl], host[out][sl])


@pytest.mark.parametrize('name', list(FORMS))
def test_fp64_mfma_many_rounds_equal_streaming_form(name, device):
  """60,001 elements: several rounds of the persistent grid and a ragged last group, against the
  streaming (DPP) form of the same chain."""
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  _, stream_kernel = catalog.load(name.replace('mfma_fp64', 'tuned'))
  out = FORMS[name]
  n = 60001
  host = make_inputs(entry, n, seed=9)
  got = gpu_run(kernel, entry, host, n, device)[out]
  ref = gpu_run(stream_kernel, entry, host, n, device)[out]
  err = normwise_error(got, ref, entry.tensors[out].get_real_volume())
  print(f'{name} N={n}: normwise rel diff to the streaming form {err:.3e}')
  assert err <= 1e-12


@pytest.mark.parametrize('tag,shape', [('mfma64w8k8acc', (8, 8)), ('mfma64w8k8end', (8, 8)),
                                       ('mfma64w4k8regs', (4, 8))])
def test_fp64_mfma_schedule_options(tag, shape, device):
  """The GEMM form's other shapes and destination placements (schedule knobs) against the
  oracle, with flags and extra offsets, on a ragged batch."""
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(f'opt_gemm_f64_{tag}')
  sch = entry.generator.get_emitted().schedule
  assert sch['kind'] == 'mfma_fp64'
  assert (sch['waves_per_block'], sch['k_block']) == shape
  n = 301
  offsets = {'B': 3, 'C': 5}
  host = make_inputs(entry, n, seed=12, offsets=offsets)
  flags = (np.arange(n) % 5 != 2).astype(np.uint32)
  ref = oracle_outputs(entry, host, n, offsets=offsets, flags=flags)['C']
  got = gpu_run(kernel, entry, host, n, device, offsets=offsets, flags=flags)['C']
  vol = entry.tensors['C'].get_real_volume()
  assert np.array_equal(got[:5], host['C'][:5])
  for e in range(n):
    sl = slice(5 + e * vol, 5 + (e + 1) * vol)
    if flags[e]:
      assert normwise_error(got[sl], ref[sl], vol) <= 1e-12
    else:
      assert np.array_equal(got[sl], host['C'][sl])
