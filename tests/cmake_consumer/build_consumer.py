"""Generates a yateto-style routine file for the C++ consumer and builds it with CMake.

    python tests/cmake_consumer/build_consumer.py <build_dir>

Generation goes through the yateto frontend the way yateto's code generator drives it:
``add_tensor`` records, an ``add_operation`` multilinear record (C = A * B), ``add_scalar``
pointwise statements (C = max(C, 0) * dt, fused into the chain's kernel), then
``generate(cpp, routine_cache)``, whose writers append the kernel and launcher to the routine
file (``kernels.cpp``) and return the prototype (``kernels.h``, with the emitted call site
wrapped in ``tf_execute``). CMake then builds
``consumer`` against ``find_package(TensorForge)``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


class _Arch:
  name, backend, typename = 'gfx950', 'hip', 'double'


class _Bounds:
  """What a yateto tensor's ``eqspp()`` answers: the box of its nonzeros (inclusive bounds)."""

  def __init__(self, shape):
    self.shape = shape

  def nnzbounds(self):
    return [(0, s - 1) for s in self.shape]


class _YatetoTensor:
  def __init__(self, name, shape):
    self._name, self._bounds = name, _Bounds(shape)

  def name(self):
    return self._name

  def eqspp(self):
    return self._bounds


class _YatetoScalar:
  datatype = None

  def __init__(self, name):
    self._name = name

  def name(self):
    return self._name


def generate(out_dir):
  sys.path.insert(0, ROOT)
  import tensorforge
  from tensorforge.generators import optree as ot
  fe = tensorforge.get_routine_generator(None)(_Arch())
  for name, shape, addressing in (('A', [56, 56], '&'), ('B', [56, 9], 'n*N+o&'),
                                  ('C', [56, 9], 'n*N+o&')):
    fe.add_tensor({'name': name, 'datatype': 'f64', 'addressing': addressing,
                   'storage': {'shape': shape, 'type': 'full'}, 'values': None,
                   'flags': {'temporary': False, 'constant': False}})
  ref = lambda n: {'name': n, 'spp': None}   # noqa: E731
  fe.add_operation({'type': 'multilinear', 'result': ref('C'), 'args': [ref('A'), ref('B')],
                    'condition': [], 'target': [[0, -1], [-1, 1]], 'permute': [[0, 1], [0, 1]],
                    'linear': {'alpha': None, 'add': False}})
  c_node, dt_node = _YatetoTensor('C', [56, 9]), _YatetoScalar('dt')
  c = ot.tensor(c_node)
  fe.add_scalar([c_node], [ot.assign(c, ot.max(c, 0.0) * ot.scalar(dt_node))], [[-1, -2]])

  os.makedirs(out_dir, exist_ok=True)
  routine_file = os.path.join(out_dir, 'kernels.cpp')
  if os.path.exists(routine_file):
    os.remove(routine_file)
  calls, prototypes = [], []

  class RoutineCache:
    def addRoutine(self, name, writer):
      prototypes.append(writer(name, routine_file))

  fe.generate(calls.append, RoutineCache())
  with open(os.path.join(out_dir, 'kernels.h'), 'w') as f:
    f.write('#pragma once\n#include <stddef.h>\n')
    f.write('\n'.join(p.strip() for p in prototypes) + '\n')
    f.write('inline void tf_execute(double* A, double* B, double* C, double dt, '
            'unsigned extraOffset_B, unsigned extraOffset_C, size_t numElements, '
            'unsigned* flags, void* streamPtr) {\n')
    for line in calls:
      f.write(f'  {line}\n')
    f.write('}\n')
  return calls


def _run(cmd, env):
  proc = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
  if proc.returncode != 0:
    raise RuntimeError(f'{" ".join(cmd)} failed:\n{proc.stdout[-4000:]}')


def build(out_dir, jobs=8):
  out_dir = os.path.abspath(out_dir)
  gen_dir = os.path.join(out_dir, 'generated')
  generate(gen_dir)
  cmake = shutil.which('cmake')
  if cmake is None:
    raise RuntimeError('cmake is not on PATH')
  env = dict(os.environ)
  env.setdefault('HIPCXX', '/opt/rocm/llvm/bin/clang++')
  _run([cmake, '-S', HERE, '-B', out_dir, '-DCMAKE_BUILD_TYPE=Release',
        f'-DTensorForge_DIR={os.path.join(ROOT, "tensorforge_amd", "share", "cmake")}',
        f'-DGENERATED_DIR={gen_dir}', '-DCMAKE_PREFIX_PATH=/opt/rocm'], env)
  _run([cmake, '--build', out_dir, '-j', str(jobs)], env)
  return os.path.join(out_dir, 'consumer')


if __name__ == '__main__':
  print(build(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, '_build')))
