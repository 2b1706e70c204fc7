"""The omptarget backend (host C++/OpenMP launchers) against the reference's own OpenMP-generated
kernels, bit for bit (CPU).

``tests/golden/omp_fixtures.{json,npz}`` hold what the reference's omptarget kernels wrote on
seeded inputs for twelve description lists (``tests/golden/make_omp_fixtures.py``). Here the
same lists go through this package's ``Context(arch='gfx950', backend='omptarget')``: the
launcher keeps the reference's name and parameters, is compiled with g++ ``-fopenmp`` (no fused
multiply-add) and run on the same inputs through its C ABI. Every output must equal the
reference's exactly, except the ``GemmDescr`` with a numeric ``alpha`` (the reference reads its
alpha back as absent and writes zeros; this package applies it, a documented deviation), which is
checked against the closed form instead. Flags and extra offsets, which the fixtures leave at
their defaults, are checked against the same launcher without them.
"""
import ctypes
import os

import numpy as np
import pytest

import omp_fixture_cases as fx

BUILD_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'omp_emitter_build', '_build')
REC = {r['label']: r for r in fx.CASES}
ZEROS_IN_REFERENCE = {'gemm_descr_alpha'}


def _generator(rec):
  from tensorforge.common.basic_types import FloatingPointType
  from tensorforge.common.context import Context
  from tensorforge.generators.generator import Generator
  fp_name, descrs = fx.descriptions(rec['label'])
  fp = {'double': FloatingPointType.DOUBLE, 'float': FloatingPointType.FLOAT}[fp_name]
  gen = Generator(descrs, Context(arch='gfx950', backend='omptarget', fp_type=fp))
  gen.generate()
  return gen


def _library(rec):
  from tensorforge_amd.runtime.build import compile_host_source
  path = os.path.join(BUILD_DIR, f'libomp_emit_{rec["label"]}.so')
  if not os.path.exists(path):
    compile_host_source(_generator(rec).get_source(), f'omp_emit_{rec["label"]}', BUILD_DIR)
  return ctypes.CDLL(path)


def _call(rec, arrays, flags=None, offsets=None):
  """Run the case's omptarget launcher on host ``arrays`` ({alias: array | [arrays] | float});
  returns the arrays after the call."""
  from tensorforge_amd.catalog import omp_arguments
  gen = _generator(rec)
  fn = getattr(_library(rec), f'launcher_{gen.get_base_name()}')
  fn.restype = None
  args, keep = omp_arguments(gen, arrays, rec['elements'], flags, offsets)
  fn(*args)
  return arrays


def _host(rec):
  gen = _generator(rec)
  return fx.host_arrays(rec, {t.alias: t for t in gen._global_tensors})


def test_backend_keeps_the_launcher_name_and_parameters():
  for rec in fx.CASES:
    gen = _generator(rec)
    assert gen.get_base_name() == rec['name']
    src = gen.get_source()
    assert f'extern "C" void {rec["launcher"]}(' in src
    assert '#pragma omp parallel for' in src


@pytest.mark.parametrize('label', [lb for lb in REC if lb not in ZEROS_IN_REFERENCE])
def test_omptarget_launcher_equals_the_reference_kernel_bitwise(label):
  rec = REC[label]
  out = _call(rec, _host(rec))
  for alias in rec['outputs']:
    got = out[alias]
    if isinstance(got, list):
      got = np.concatenate(got)
    np.testing.assert_array_equal(got, fx.expected(rec, alias), err_msg=f'{label} {alias}')


def test_alpha_is_applied_where_the_reference_writes_zeros():
  rec = REC['gemm_descr_alpha']
  host = _host(rec)
  n = rec['elements']
  a = host['Na'].reshape(7, 10).T
  b = host['Nb'].reshape(n, 5, 7).transpose(0, 2, 1)
  want = 2.5 * np.einsum('ik,ekj->eji', a, b).reshape(-1)
  out = _call(rec, host)
  assert not fx.expected(rec, 'Nc').any()
  np.testing.assert_allclose(out['Nc'], want, rtol=1e-13, atol=1e-13)


def test_flags_and_extra_offsets():
  rec = REC['ader_volume_o6_sparse']
  base = _host(rec)
  n, vol = rec['elements'], 504
  plain = _call(rec, {k: (v.copy() if hasattr(v, 'copy') else v) for k, v in base.items()})
  shifted = {}
  offsets = {}
  for p in rec['params']:
    v = base[p['alias']]
    if p['addressing'] == 'STRIDED':
      offsets[p['alias']] = 3
      shifted[p['alias']] = np.concatenate([np.full(3, 7.0, dtype=v.dtype), v])
    else:
      shifted[p['alias']] = v
  flags = (np.arange(n) % 3 != 1).astype(np.uint32)
  out = _call(rec, shifted, flags=flags, offsets=offsets)
  q = out['Qp']
  assert np.all(q[:3] == 7.0)
  for e in range(n):
    sl = slice(3 + e * vol, 3 + (e + 1) * vol)
    ref = plain['Qp'][e * vol:(e + 1) * vol] if flags[e] else base['Qp'][e * vol:(e + 1) * vol]
    np.testing.assert_array_equal(q[sl], ref)


def test_null_stream_throws_invalid_argument(tmp_path):
  """A null ``streamPtr`` throws ``std::invalid_argument`` before any work, as the reference's
  omptarget launcher does; a C++ caller catches it."""
  import subprocess
  from tensorforge_amd.runtime.build import host_cxx
  rec = REC['ader_volume_o6_sparse']
  gen = _generator(rec)
  lib = os.path.join(BUILD_DIR, f'libomp_emit_{rec["label"]}.so')
  _library(rec)
  proto = gen._generate_launcher_proto(with_defaults=False)
  n_params = proto[proto.index('(') + 1:proto.rindex(')')].count(',') + 1
  call = f'launcher_{gen.get_base_name()}({", ".join(["0"] * n_params)})'
  src = tmp_path / 'caller.cpp'
  src.write_text('#include <cstdio>\n#include <cstddef>\n#include <stdexcept>\n'
                 f'extern "C" {proto};\n'
                 'int main() {\n'
                 f'  try {{ {call}; }} catch (const std::invalid_argument& e) {{\n'
                 '    std::printf("caught: %s\\n", e.what()); return 0; }\n'
                 '  return 1;\n}\n')
  exe = tmp_path / 'caller'
  subprocess.run([host_cxx(), '-std=c++17', str(src), lib, f'-Wl,-rpath,{BUILD_DIR}', '-o',
                  str(exe)], check=True)
  res = subprocess.run([str(exe)], capture_output=True, text=True)
  assert res.returncode == 0 and 'stream may not be null' in res.stdout
