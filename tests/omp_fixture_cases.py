"""Shared plumbing of the tests that read the OpenMP golden fixtures
(``tests/golden/omp_fixtures.{json,npz}``, written by ``tests/golden/make_omp_fixtures.py`` from
the reference's own omptarget-generated kernels): the cases, their inputs regenerated from the
seed, and this package's generator for the same description list."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLDEN_DIR = os.path.join(HERE, 'golden')
if GOLDEN_DIR not in sys.path:
  sys.path.insert(0, GOLDEN_DIR)

import name_cases  # noqa: E402
from omp_inputs import fixture_inputs  # noqa: E402

with open(os.path.join(GOLDEN_DIR, 'omp_fixtures.json'), encoding='utf-8') as _f:
  RECORDS = json.load(_f)
FIXTURES = np.load(os.path.join(GOLDEN_DIR, 'omp_fixtures.npz'), allow_pickle=False)
CASES = [r for r in RECORDS if 'error' not in r]
SKIPPED = [r for r in RECORDS if 'error' in r]


def descriptions(label):
  for lab, fp_name, descrs in name_cases.cases() + name_cases.extra_cases():
    if lab == label:
      return fp_name, descrs
  raise KeyError(label)


def dtype_of(rec):
  return np.float64 if rec['fp_type'] == 'double' else np.float32


def inputs(rec):
  return fixture_inputs(rec['params'], rec['elements'], rec['seed'], dtype_of(rec))


def expected(rec, alias):
  return FIXTURES[f'{rec["label"]}__{alias}']


def generator(rec, emitter='auto'):
  """This package's generator for the case (generated), and ``{alias: Tensor}``."""
  import tensorforge
  assert os.path.abspath(tensorforge.__file__).startswith(REPO)
  from tensorforge.common.basic_types import FloatingPointType
  from tensorforge.common.context import Context
  from tensorforge.generators.generator import Generator
  fp_name, descrs = descriptions(rec['label'])
  fp = {'double': FloatingPointType.DOUBLE, 'float': FloatingPointType.FLOAT}[fp_name]
  gen = Generator(descrs, Context(arch='gfx950', backend='hip', fp_type=fp), emitter=emitter)
  gen.generate()
  tensors = {t.alias: t for t in gen._global_tensors}
  return gen, descrs, tensors


def host_arrays(rec, tensors):
  """Fixture inputs in the launcher layout: PTR_BASED operands as per-element arrays."""
  out = {}
  raw = inputs(rec)
  for p in rec['params']:
    v = raw[p['alias']]
    if p['addressing'] == 'PTR_BASED':
      v = [v[e * p['volume']:(e + 1) * p['volume']].copy() for e in range(rec['elements'])]
    out[p['alias']] = v
  return out
