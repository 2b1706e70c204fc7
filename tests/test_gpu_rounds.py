"""Persistent kernels over many element groups per workgroup.

The streaming, column-per-lane and matrix-core kernels size their grid to the resident
workgroups and loop over element groups, carrying prefetched rows from one group to the next.
The parity tests' batches fit in one round; here every tuned launcher runs 60,001 elements
(at least two rounds on MI355X, a partial last group) and is compared on every element with the
plain launcher of the same chain, which has no cross-group state (itself checked against the
oracle). Normwise per element: 1e-12 FP64, 1e-5 FP32, FP32 MFMA and split BF16x3.
"""
import pytest

from tensorforge_amd import catalog
from tensorforge_amd.common.basic_types import FloatingPointType

from harness import gpu_run, key_of, make_inputs, normwise_error, outputs_of

pytestmark = pytest.mark.gpu

N = 60001
PAIRS = [(t, t.rsplit('_', 1)[0] + '_plain', 1e-5 if '_f32_' in t else 1e-12)
         for t in catalog.names() if t.endswith('_tuned') and not t.startswith('t_')]
PAIRS += [('ader7_f64_mfma_fp32', 'ader7_f64_plain', 1e-5),
          ('ader7_f64_mfma_bf16x3', 'ader7_f64_plain', 1e-5)]
PAIRS += [(t, t[:-len('_auto')] + '_plain',
           1e-5 if catalog.get(t).fp_type == FloatingPointType.FLOAT else 1e-12)
          for t in catalog.names() if t.startswith('t_') and t.endswith('_auto')]
# the streaming emitter's schedule options (catalog opt_<chain>_<option>)
PAIRS += [(t, t[len('opt_'):t.rindex('_')] + '_plain', 1e-12)
          for t in catalog.names() if t.startswith('opt_')]


@pytest.mark.parametrize('tuned,plain,tol', PAIRS)
def test_many_rounds_match_plain(tuned, plain, tol, device):
  et, kt = catalog.load(tuned)
  ep, kp = catalog.load(plain)
  print(f'{tuned} ({et.generator.get_emitted().emitter}) vs {plain}, N={N}', flush=True)
  host = make_inputs(ep, N, seed=17)
  a = gpu_run(kp, ep, host, N, device)
  b = gpu_run(kt, et, host, N, device)
  for t in ep.generator._global_tensors:
    key = key_of(t)
    if key not in outputs_of(ep):
      continue
    err = normwise_error(b[key], a[key], t.get_real_volume())
    print(f'{tuned} {key}: normwise rel diff {err:.3e} (tol {tol})', flush=True)
    assert err <= tol


@pytest.mark.parametrize('tuned,plain', [('ader_f64_tuned', 'ader_f64_plain'),
                                         ('gemm_f64_tuned', 'gemm_f64_plain'),
                                         ('flux_neigh_mod_f64_tuned', 'flux_neigh_mod_f64_plain'),
                                         # odd offsets: the direct-to-LDS loads' fallback kernel
                                         ('opt_ader_f64_dma', 'ader_f64_plain')])
def test_many_rounds_with_flags_and_offsets(tuned, plain, device):
  """Skipped elements and extra offsets across rounds: untouched entries stay untouched."""
  import numpy as np
  et, kt = catalog.load(tuned)
  ep, kp = catalog.load(plain)
  keys = [key_of(t) for t in ep.generator._global_tensors if t.addressing.name == 'STRIDED']
  offsets = {k: 3 + 2 * i for i, k in enumerate(keys)}
  flags = (np.arange(N) % 5 != 3).astype(np.uint32)
  print(f'{tuned} vs {plain}, N={N}, offsets {offsets}, every fifth element skipped', flush=True)
  host = make_inputs(ep, N, seed=23, offsets=offsets)
  a = gpu_run(kp, ep, host, N, device, offsets=offsets, flags=flags)
  b = gpu_run(kt, et, host, N, device, offsets=offsets, flags=flags)
  for t in ep.generator._global_tensors:
    key = key_of(t)
    if key not in outputs_of(ep):
      continue
    off = offsets.get(key, 0)
    np.testing.assert_array_equal(b[key][:off], host[key][:off])
    vol = t.get_real_volume()
    skipped = np.where(flags == 0)[0]
    got = b[key][off:off + N * vol].reshape(N, vol)
    orig = host[key][off:off + N * vol].reshape(N, vol)
    np.testing.assert_array_equal(got[skipped], orig[skipped])
    err = normwise_error(b[key], a[key], vol, off)
    print(f'{tuned} {key}: normwise rel diff {err:.3e}', flush=True)
    assert err <= 1e-12
