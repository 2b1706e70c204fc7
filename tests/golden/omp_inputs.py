"""Input recipe of the OpenMP golden fixtures (``omp_fixtures.json`` / ``omp_fixtures.npz``),
shared by the script that wrote them (``make_omp_fixtures.py``) and the tests that read them:
the inputs are regenerated from the case's seed, only the outputs are stored.

``params`` is the case's list of global operands in launcher-parameter order, each
``{'alias', 'addressing', 'volume', 'count'}``: a STRIDED or PTR_BASED operand holds ``volume``
entries per element (its full storage, also for a packed sparse tensor), a shared (NONE) one
``count`` entries (its stored nonzeros when sparse), a SCALAR one a single value.
"""
import numpy as np


def fixture_inputs(params, n, seed, dtype):
  """``{alias: array | float}``: uniform [-1, 1] entries (SCALAR: uniform [0.5, 2]), drawn in
  parameter order from ``numpy.random.default_rng(seed)`` in FP64 and rounded to ``dtype``."""
  rng = np.random.default_rng(seed)
  out = {}
  for p in params:
    if p['addressing'] == 'SCALAR':
      out[p['alias']] = float(np.asarray(rng.uniform(0.5, 2.0)).astype(dtype))
    elif p['addressing'] == 'NONE':
      out[p['alias']] = rng.uniform(-1, 1, p['count']).astype(dtype)
    else:
      out[p['alias']] = rng.uniform(-1, 1, n * p['volume']).astype(dtype)
  return out
