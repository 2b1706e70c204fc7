"""The C-ABI objects load and export every symbol the headers declare (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from tensorforge_amd import catalog
from tensorforge_amd.runtime import build as rt_build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
  text = open(os.path.join(ROOT, 'include', header)).read()
  text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
  text = '\n'.join(l for l in text.splitlines() if not l.lstrip().startswith('#'))
  return re.findall(r'\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;', text)


def _exports(path):
  out = subprocess.run(['nm', '-D', '--defined-only', path], stdout=subprocess.PIPE, text=True,
                       check=True).stdout
  return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_runtime_exports_header_symbols():
  path = rt_build.runtime_path()
  assert os.path.exists(path), 'build() first'
  exported = _exports(path)
  declared = _declared('tensorforge_amd.h') + _declared('tensorforge_mlp.h')
  assert declared, 'no declarations parsed'
  missing = [s for s in declared if s not in exported]
  assert not missing, missing
  ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def test_catalog_objects_export_declared_launchers():
  declared = set(_declared('tensorforge_catalog.h'))
  assert declared
  exported = set()
  for name in catalog.names():
    path = catalog.so_path(name)
    assert os.path.exists(path), f'{path} not built'
    exported |= _exports(path)
  for name in catalog.group_names():
    for part in ('tuned', 'plain'):
      path = catalog.group_so_path(name, part)
      assert os.path.exists(path), f'{path} not built'
      exported |= _exports(path)
  missing = sorted(declared - exported)
  assert not missing, missing


def test_each_entry_exports_its_launcher():
  for name in catalog.names():
    entry = catalog.get(name)
    entry.generator.register()
    assert f'launcher_{entry.generator.get_base_name()}' in _exports(catalog.so_path(name)), name


def test_header_matches_current_generators():
  from tensorforge_amd import catalog_header
  assert open(catalog_header.HEADER_PATH).read() == catalog_header.render(), \
    'include/tensorforge_catalog.h is stale: run build()'


def test_loader_binds_without_gpu_calls():
  entry = catalog.get('gemm_f64_plain')
  entry.generator.generate()
  from tensorforge_amd.runtime.loader import load_kernels
  k = load_kernels(catalog.so_path('gemm_f64_plain'), [entry.generator])[0]
  assert [t for t, _ in k.signature][-3:] == ['size_t', 'unsigned*', 'void*']
  with pytest.raises(TypeError):
    k(1, 2)


def test_shared_launcher_names_resolve_per_object():
  """Chains that differ only in floating type or sparsity share a launcher name; with its
  TF_CATALOG_<ENTRY> macro defined the header declares exactly that object's prototype."""
  from tensorforge_amd import catalog_header
  header = os.path.join(ROOT, 'include', 'tensorforge_catalog.h')
  for name in ('ader_f64_tuned', 'ader_f32_tuned', 'gemm_f64_tuned', 'gemm_f32_plain',
               'pw_update_f64', 'pw_update_f32'):
    g = catalog.get(name).generator
    g.register()
    proto = g._generate_launcher_proto(with_defaults=False).replace('unsigned* flags ',
                                                                    'unsigned* flags')
    out = subprocess.run(['gcc', '-E', '-P', '-x', 'c', f'-D{catalog_header.macro(name)}', header],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    decls = [l for l in out.splitlines() if f'launcher_{g.get_base_name()}(' in l]
    assert decls == [proto + ';'], (name, decls)
  r = subprocess.run(['gcc', '-fsyntax-only', '-x', 'c', '-DTF_CATALOG_ADER_F64_TUNED', header],
                     capture_output=True, text=True)
  assert r.returncode == 0, r.stderr
