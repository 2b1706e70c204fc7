"""The chains either side of the volume chain, with the reference benchmarks' shapes, tuned and
plain launchers against the FP64 oracle:

* local and neighbour flux integrals: 64-row padded storage, bounding boxes 56x9, 56x21, 21x56,
  21x21, in both operand orderings;
* the time-derivative step ``D += C (A B) + F (A E) + H (A G)`` with 48x53/54/55 shared matrices
  over 64-row arrays, aligned (one streamed array) and not aligned (three arrays read from row 1,
  35 of 48 destination rows)."""
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


DERIV = [f'deriv{v}_f64_{em}' for v in ('', '_na') for em in ('tuned', 'plain')]


@pytest.mark.parametrize('name', DERIV)
def test_derivative_chain_matches_oracle(name, device):
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  n = 300
  host = make_inputs(entry, n, seed=15)
  ref = oracle_outputs(entry, host, n)
  got = gpu_run(kernel, entry, host, n, device)
  print(f'{name}: {entry.generator.get_emitted().emitter} '
        f'{entry.generator.get_emitted().schedule.get("kind")}', flush=True)
  err = normwise_error(got['D'], ref['D'], 48 * 9)
  print(f'{name} D: {err:.2e}', flush=True)
  assert err <= TOL[entry.fp_type]


@pytest.mark.parametrize('name', ['deriv_f64_tuned', 'deriv_na_f64_tuned'])
def test_derivative_flags_offsets_and_window(name, device):
  """Skipped elements keep their values; rows 35..47 of the not-aligned destination are
  outside its bounding box and never written."""
  from tensorforge_amd import catalog
  entry, kernel = catalog.load(name)
  n = 41
  offsets = {'D': 5, 'S2': 3}
  host = make_inputs(entry, n, seed=16, offsets=offsets)
  flags = (np.arange(n) % 3 != 1).astype(np.uint32)
  ref = oracle_outputs(entry, host, n, offsets=offsets, flags=flags)['D']
  got = gpu_run(kernel, entry, host, n, device, offsets=offsets, flags=flags)['D']
  err = normwise_error(got, ref, 48 * 9, offset=5)
  print(f'{name} D with flags/offsets: {err:.2e}', flush=True)
  assert err <= 1e-12
  g = got[5:].reshape(n, 9, 48)
  h = host['D'][5:].reshape(n, 9, 48)
  assert np.array_equal(g[flags == 0], h[flags == 0])
  rows = entry.generator.descr_list[-1].dest.bbox.upper()[0]
  assert np.array_equal(g[:, :, rows:], h[:, :, rows:])
