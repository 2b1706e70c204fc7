"""Build support for tests/test_gpu_omp_fixtures.py: every OpenMP fixture case
(tests/golden/omp_fixtures.json) compiled with the plain and with the default ('auto') emitter,
one shared object per case and emitter (cases of one description but different floating types
or sparsity share a launcher name), under tests/omp_fixture_build/_build.

    python tests/omp_fixture_build.py
"""
import concurrent.futures
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
  if p not in sys.path:
    sys.path.insert(0, p)

BUILD_DIR = os.path.join(HERE, 'omp_fixture_build', '_build')
EMITTERS = ('plain', 'auto')


def so_path(label, emitter, directory=BUILD_DIR):
  return os.path.join(directory, f'libomp_{label}_{emitter}.so')


def build(directory=BUILD_DIR, jobs=8):
  import omp_fixture_cases as fx
  from tensorforge_amd.emit.launcher import translation_unit
  from tensorforge_amd.runtime.build import build_runtime, compile_source, runtime_path
  if not os.path.exists(runtime_path()):
    build_runtime()
  work = []
  for rec in fx.CASES:
    for emitter in EMITTERS:
      gen, _, _ = fx.generator(rec, emitter)
      work.append((translation_unit([gen]), f'omp_{rec["label"]}_{emitter}'))
  with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
    list(pool.map(lambda w: compile_source(w[0], w[1], directory, keep_source=False), work))
  return len(work)


if __name__ == '__main__':
  print(build(), 'objects in', BUILD_DIR)
