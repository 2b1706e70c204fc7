"""The omptarget backend against the FP64 oracle on every plain catalog chain (CPU).

Each chain's descriptions generate a host C++/OpenMP launcher (``Context(backend='omptarget')``,
``tensorforge_amd/emit/omp.py``), built by ``tests/omp_catalog_build.py``; it runs on the same
seeded inputs as the GPU tests, through its C ABI, and every output must be within the stated
tolerance of the oracle (normwise, 1e-12 FP64 / 1e-5 FP32). The chains cover sparse and dense
shared operands, transposes with alpha and beta, bounding boxes, 3-index contractions,
pointer-based and strided operands, temporaries and the flux and derivative chains. One case
also runs with flags and extra offsets: skipped elements and the region before the offset stay
untouched.
"""
import ctypes
import os

import numpy as np
import pytest

import omp_catalog_build as ob
from harness import TOL, key_of, make_inputs, normwise_error, oracle_outputs, outputs_of

_LIBS = {}


def _launcher(name):
  from tensorforge_amd import catalog
  if name not in _LIBS:
    entry = catalog.get(name)
    if entry.generator.get_kernel() is None:     # the inputs follow its launcher's operands
      entry.generator.generate()
    gen = ob.generator(entry)
    path = ob.so_path(name)
    if not os.path.exists(path):
      ob.build_one(name)
    fn = getattr(ctypes.CDLL(path), f'launcher_{gen.get_base_name()}')
    fn.restype = None
    _LIBS[name] = (entry, gen, fn)
  return _LIBS[name]


def omp_run(name, host, n, offsets=None, flags=None):
  from tensorforge_amd.catalog import omp_arguments
  entry, gen, fn = _launcher(name)
  arrays = {k: ([x.copy() for x in v] if isinstance(v, list) else
                (v.copy() if hasattr(v, 'copy') else v)) for k, v in host.items()}
  args, keep = omp_arguments(gen, arrays, n, flags=flags, offsets=offsets)
  fn(*args)
  del keep
  return arrays


def _check(name, n, seed=0, offsets=None, flags=None):
  entry = _launcher(name)[0]
  host = make_inputs(entry, n, seed=seed, offsets=offsets)
  ref = oracle_outputs(entry, host, n, offsets=offsets, flags=flags)
  got = omp_run(name, host, n, offsets=offsets, flags=flags)
  tol = TOL[entry.fp_type]
  checked = 0
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
      np.testing.assert_array_equal(got[key][:off], host[key][:off])
    print(f'{name} {key} N={n} normwise rel err {err:.3e} (tol {tol})')
    assert err <= tol, f'{name}: {key} normwise error {err:.3e} > {tol}'
    checked += 1
  assert checked


@pytest.mark.parametrize('name', ob.names())
def test_omptarget_launcher_matches_oracle(name):
  _check(name, 37)


def test_omptarget_flags_and_offsets():
  entry = _launcher('t_four_matrices_plain')[0]
  offsets = {key_of(t): 5 for t in entry.generator._global_tensors
             if t.addressing.name in ('STRIDED', 'PTR_BASED')}
  flags = (np.arange(41) % 4 != 2).astype(np.uint32)
  _check('t_four_matrices_plain', 41, seed=3, offsets=offsets, flags=flags)
