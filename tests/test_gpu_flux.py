"""Local and neighbour flux-integral chains (the steps either side of the volume chain),
with the reference benchmark's shapes: 64-row padded storage, bounding boxes 56x9, 56x21,
21x56, 21x21, in both operand orderings. Tuned and plain launchers against the FP64 oracle."""
import numpy as np
import pytest

from harness import TOL, gpu_run, make_inputs, normwise_error, oracle_outputs, outputs_of

pytestmark = pytest.mark.gpu

NAMES = [f'flux_{kind}_{order}_f64_{em}' for kind in ('local', 'neigh') for order in ('mod', 'orig')
         for em in ('tuned', 'plain')]


@pytest.mark.parametrize('name', NAMES)
def test_flux_chain_matches_oracle(name, device):
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  n = 300
  host = make_inputs(entry, n, seed=13)
  ref = oracle_outputs(entry, host, n)
  got = gpu_run(kernel, entry, host, n, device)
  for key in outputs_of(entry):
    t = next(t for t in entry.generator._global_tensors if (t.alias or t.name) == key)
    err = normwise_error(got[key], ref[key], t.get_real_volume())
    print(f'{name} {key}: {err:.2e}')
    assert err <= TOL[entry.fp_type]


@pytest.mark.parametrize('kind', ['local', 'neigh'])
def test_flux_flags_offsets_and_padding(kind, device):
  """Rows 56..63 of the padded destination are outside the bounding box: never written."""
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(f'flux_{kind}_mod_f64_tuned')
  n = 23
  dest = 'i_surf' if kind == 'local' else 'dofs'
  src = 'd_k' if kind == 'local' else 'i_dofs'
  offsets = {dest: 3, src: 2, 'a_plus': 1}
  host = make_inputs(entry, n, seed=14, offsets=offsets)
  flags = (np.arange(n) % 4 != 2).astype(np.uint32)
  ref = oracle_outputs(entry, host, n, offsets=offsets, flags=flags)[dest]
  got = gpu_run(kernel, entry, host, n, device, offsets=offsets, flags=flags)[dest]
  assert normwise_error(got, ref, 64 * 9, offset=3) <= 1e-12
  g = got[3:].reshape(n, 9, 64)
  h = host[dest][3:].reshape(n, 9, 64)
  assert np.array_equal(g[:, :, 56:], h[:, :, 56:])
  assert np.array_equal(g[flags == 0], h[flags == 0])
