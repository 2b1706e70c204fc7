"""Write ``omp_fixtures.json`` and ``omp_fixtures.npz``: outputs of the reference's own OpenMP
("omptarget") generated kernels on seeded inputs, for every description list of
``name_cases.cases()`` and ``name_cases.extra_cases()`` that the reference generates.

Run in a checkout that holds the reference next to this repository:

    python tests/golden/make_omp_fixtures.py

For each case it imports the reference's ``tensorforge`` from ``/root/reference``, builds the
description list with it, generates with ``Context(arch='gfx950', backend='omptarget')``
(reference ``common/vm/lexic/target_lexic.py:39-205``, ``generators/generator.py:102-199``), writes
the generated kernel and launcher with a small driver into a temporary directory, compiles them with
``g++ -O3 -march=x86-64-v2 -fopenmp`` together with the reference's
``include/tensorforge_aux_target.cpp`` (the flags of ``oracle/Makefile``: no FMA contraction on the
host), runs the driver and keeps only the numbers. The driver calls the launcher once per block of
``mults`` elements (GCC's host fallback runs one team, and the generated code takes the block from
the team number), with a non-null stream object and ``flags = nullptr``, then
``tensorforge::synchDevice``. Nothing of the reference or of its generated code is written into
this repository. Inputs follow ``omp_inputs.fixture_inputs`` (regenerated from the seed by the
tests); the fixture holds the outputs, as the reference computed them, per case.
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = '/root/reference'
BLOCKS = 2            # launcher calls (blocks of `mults` elements) per case
SEED = 2024


def _params(header, tensors_by_name):
  """Launcher parameters from the reference's header prototype."""
  inner = header[header.index('(') + 1:header.rindex(')')]
  out = []
  for decl in [d.strip() for d in inner.split(',')]:
    decl = decl.split('=')[0].strip()
    ctype, name = decl.rsplit(' ', 1)
    out.append((ctype.strip(), name))
  return out


def _driver(case, kernel_src, launcher_src, params, tensors_by_name, dtype, mults, n):
  """C++ source of the driver: reads the inputs, launches per block, writes the outputs."""
  real = 'double' if dtype == np.float64 else 'float'
  lines = ['#include <cstdio>', '#include <cstdlib>', '#include <cstring>', '#include <stdexcept>',
           '#include <cmath>', '#include <string>', '#include <vector>', '#include <omp.h>',
           '#include "tensorforge_aux.h"', '', kernel_src, '', launcher_src, '',
           'int main(int argc, char** argv) {',
           '  if (argc != 3) return 2;',
           '  FILE* fin = std::fopen(argv[1], "rb");',
           '  if (!fin) return 3;']
  args = []
  outputs = []
  for ctype, name in params:
    if name in ('numElements', 'flags', 'streamPtr'):
      continue
    if name.endswith('_extraOffset'):
      continue
    t = tensors_by_name[name]
    addr = t['addressing']
    if addr == 'SCALAR':
      lines.append(f'  {real} {name} = 0; if (std::fread(&{name}, sizeof({real}), 1, fin) != 1) return 4;')
      continue
    count = t['count'] if addr == 'NONE' else n * t['volume']
    lines.append(f'  std::vector<{real}> {name}_buf({count});')
    lines.append(f'  if (std::fread({name}_buf.data(), sizeof({real}), {count}, fin) != {count}) return 4;')
    if addr == 'PTR_BASED':
      lines.append(f'  std::vector<{real}*> {name}_tab({n});')
      lines.append(f'  for (int e = 0; e < {n}; ++e) {name}_tab[e] = {name}_buf.data() + '
                   f'(size_t)e * {t["volume"]};')
    if not t['const']:
      outputs.append((name, count))
  lines.append('  std::fclose(fin);')
  lines.append('  int stream_obj = 0;')
  lines.append(f'  for (int blk = 0; blk < {n // mults}; ++blk) {{')
  call = []
  for ctype, name in params:
    if name == 'numElements':
      call.append(str(mults))
    elif name == 'flags':
      call.append('nullptr')
    elif name == 'streamPtr':
      call.append('&stream_obj')
    elif name.endswith('_extraOffset'):
      call.append('0u')
    else:
      t = tensors_by_name[name]
      addr = t['addressing']
      if addr == 'SCALAR':
        call.append(name)
      elif addr == 'NONE':
        call.append(f'const_cast<{ctype}>({name}_buf.data())')
      elif addr == 'STRIDED':
        call.append(f'const_cast<{ctype}>({name}_buf.data() + (size_t)blk * {mults} * {t["volume"]})')
      else:
        call.append(f'const_cast<{ctype}>({name}_tab.data() + (size_t)blk * {mults})')
  lines.append(f'    {case["launcher"]}({", ".join(call)});')
  lines.append('  }')
  lines.append('  tensorforge::synchDevice(&stream_obj);')
  lines.append('  FILE* fout = std::fopen(argv[2], "wb");')
  lines.append('  if (!fout) return 5;')
  for name, count in outputs:
    lines.append(f'  std::fwrite({name}_buf.data(), sizeof({real}), {count}, fout);')
  lines.append('  std::fclose(fout);')
  lines.append('  return 0;')
  lines.append('}')
  return '\n'.join(lines) + '\n', outputs


def main():
  sys.path.insert(0, HERE)
  sys.path.insert(0, REFERENCE)
  import tensorforge
  assert os.path.abspath(tensorforge.__file__).startswith(REFERENCE), tensorforge.__file__
  from tensorforge.common.basic_types import Addressing, DataFlowDirection, FloatingPointType
  from tensorforge.common.context import Context
  from tensorforge.generators.generator import Generator
  import name_cases
  from omp_inputs import fixture_inputs

  fp = {'double': FloatingPointType.DOUBLE, 'float': FloatingPointType.FLOAT}
  work = tempfile.mkdtemp(prefix='tf_omp_')
  records, arrays = [], {}
  try:
    for index, (label, fp_name, descrs) in enumerate(name_cases.cases() + name_cases.extra_cases()):
      dtype = np.float64 if fp_name == 'double' else np.float32
      rec = {'label': label, 'fp_type': fp_name}
      try:
        ctx = Context(arch='gfx950', backend='omptarget', fp_type=fp[fp_name])
        gen = Generator(descrs, ctx)
        gen.generate()
      except Exception as exc:
        rec['error'] = f'generation: {type(exc).__name__}: {exc}'
        records.append(rec)
        print(label, rec['error'], flush=True)
        continue
      launcher = gen.get_launcher()
      block = re.search(r'size_t block\[3\] = \{\s*(\d+),\s*(\d+),\s*(\d+)\s*\}', launcher)
      mults = int(block.group(2))
      n = BLOCKS * mults
      tensors = {}
      for d in descrs:
        for sub in [d.dest] + list(d.ops):
          t = getattr(sub, 'tensor', sub)
          if t.name is not None and not t.is_tmp:
            tensors[t.name] = {
              'alias': t.alias, 'addressing': t.addressing.name,
              'volume': int(t.get_real_volume()), 'count': int(t.memory()),
              'const': t.direction == DataFlowDirection.SOURCE}
      header = gen.get_header()
      params = _params(header, tensors)
      order = [tensors[nm] for _, nm in params if nm in tensors]
      rec.update({'name': gen.get_base_name(), 'launcher': f'launcher_{gen.get_base_name()}',
                  'block': [int(block.group(i)) for i in (1, 2, 3)], 'mults': mults,
                  'elements': n, 'seed': SEED + index,
                  'params': [{k: v for k, v in t.items() if k != 'const'} for t in order],
                  'outputs': [t['alias'] for t in order if not t['const']
                              and t['addressing'] != 'SCALAR']})
      src, outputs = _driver(rec, gen.get_kernel(), launcher, params, tensors, dtype, mults, n)
      case_dir = os.path.join(work, label)
      os.makedirs(case_dir)
      with open(os.path.join(case_dir, 'driver.cpp'), 'w') as f:
        f.write(src)
      exe = os.path.join(case_dir, 'driver')
      cmd = ['g++', '-O3', '-march=x86-64-v2', '-fopenmp', '-std=c++17',
             f'-I{REFERENCE}/tensorforge/include', os.path.join(case_dir, 'driver.cpp'),
             f'{REFERENCE}/tensorforge/include/tensorforge_aux_target.cpp', '-o', exe]
      r = subprocess.run(cmd, capture_output=True, text=True)
      if r.returncode != 0:
        rec['error'] = 'compilation: ' + r.stderr.strip().splitlines()[-1][:300]
        records.append(rec)
        print(label, rec['error'], flush=True)
        continue
      inputs = fixture_inputs(rec['params'], n, rec['seed'], dtype)
      with open(os.path.join(case_dir, 'in.bin'), 'wb') as f:
        for t in order:
          v = inputs[t['alias']]
          np.asarray(v, dtype=dtype).tofile(f)
      env = dict(os.environ, OMP_NUM_THREADS='4')
      r = subprocess.run([exe, os.path.join(case_dir, 'in.bin'), os.path.join(case_dir, 'out.bin')],
                         capture_output=True, text=True, env=env, timeout=600)
      if r.returncode != 0:
        rec['error'] = f'run: exit {r.returncode} {r.stderr[-300:]}'
        records.append(rec)
        print(label, rec['error'], flush=True)
        continue
      raw = np.fromfile(os.path.join(case_dir, 'out.bin'), dtype=dtype)
      pos = 0
      for (name, count) in outputs:
        alias = tensors[name]['alias']
        arrays[f'{label}__{alias}'] = raw[pos:pos + count].copy()
        pos += count
      assert pos == raw.size
      records.append(rec)
      print(label, rec['name'], f'{n} elements', 'ok', flush=True)
  finally:
    shutil.rmtree(work, ignore_errors=True)
  with open(os.path.join(HERE, 'omp_fixtures.json'), 'w', encoding='utf-8') as f:
    json.dump(records, f, indent=1)
    f.write('\n')
  np.savez_compressed(os.path.join(HERE, 'omp_fixtures.npz'), **arrays)


if __name__ == '__main__':
  main()
