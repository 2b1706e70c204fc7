"""Build support for tests/test_gpu_yateto_gateway.py: the kernels this package's yateto
frontend generates through the legacy gateway for every case of tests/golden/yateto_cases.py,
compiled one shared object per case under tests/yateto_gateway_build/_build.

    python tests/yateto_gateway_build.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, os.path.join(HERE, 'golden')):
  if p not in sys.path:
    sys.path.insert(0, p)

BUILD_DIR = os.path.join(HERE, 'yateto_gateway_build', '_build')


def so_path(case, directory=BUILD_DIR):
  return os.path.join(directory, f'libyateto_{case}.so')


def generators(case):
  """(frontend, [generators of the case's routines])."""
  import tensorforge
  import yateto_cases as yc
  fe, _, cache, _ = yc.run(tensorforge, case)
  return fe, [w._generator for _, w in sorted(cache.routines.items())]


def build(directory=BUILD_DIR):
  import yateto_cases as yc
  from tensorforge_amd.emit.launcher import translation_unit
  from tensorforge_amd.runtime.build import build_runtime, compile_source, runtime_path
  if not os.path.exists(runtime_path()):
    build_runtime()
  for case in yc.CASES:
    _, gens = generators(case)
    compile_source(translation_unit(gens), f'yateto_{case}', directory, keep_source=False)
  return len(yc.CASES)


if __name__ == '__main__':
  print(build(), 'objects in', BUILD_DIR)
