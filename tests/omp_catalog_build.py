"""Build support for tests/test_omp_catalog.py: the omptarget (host C++/OpenMP) launcher of every
plain catalog chain, generated with ``Context(arch='gfx950', backend='omptarget')`` from the
chain's descriptions and compiled with g++ under tests/omp_catalog_build/_build.

    python tests/omp_catalog_build.py
"""
import concurrent.futures
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
  if p not in sys.path:
    sys.path.insert(0, p)

BUILD_DIR = os.path.join(HERE, 'omp_catalog_build', '_build')


def names():
  from tensorforge_amd import catalog
  return [n for n in catalog.names() if n.endswith('_plain')]


def generator(entry):
  from tensorforge_amd.common.context import Context
  from tensorforge_amd.generators.generator import Generator
  gen = Generator(entry.descrs, Context('gfx950', 'omptarget', entry.fp_type))
  gen.generate()
  return gen


def so_path(name, directory=BUILD_DIR):
  return os.path.join(directory, f'libomp_{name}.so')


def build_one(name, directory=BUILD_DIR):
  from tensorforge_amd import catalog
  from tensorforge_amd.runtime.build import compile_host_source
  return compile_host_source(generator(catalog.get(name)).get_source(), f'omp_{name}', directory)


def build(directory=BUILD_DIR, jobs=8):
  todo = names()
  with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
    list(pool.map(lambda n: build_one(n, directory), todo))
  return len(todo)


if __name__ == '__main__':
  print(f'{build()} omptarget launchers in {BUILD_DIR}')
