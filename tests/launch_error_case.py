"""Build support for the launch-error regression test (tests/test_gpu_launch_errors.py).

Writes ``tests/launch_error_build/_build/liblaunch_error.so``: the single-GEMM chain's plain
launcher, renamed ``launcher_kernel_launch_error_probe``, whose dynamic LDS request exceeds what a
gfx950 workgroup may hold (160 KB). Its attribute call is refused by the HIP runtime on the host,
so the launcher records the error and returns without launching anything. The test then calls a
valid launcher in the same process, which must succeed: a refused call must not leave HIP's
last error set for the next launcher to read back as its own.

    python tests/launch_error_case.py [out_dir]
"""
import dataclasses
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)

BUILD_DIR = os.path.join(ROOT, 'tests', 'launch_error_build', '_build')
PROBE_NAME = 'kernel_launch_error_probe'
# more dynamic LDS than one gfx950 workgroup can address
OVER_LIMIT_LDS = 256 * 1024


def generator():
  """The probe's generator: the single-GEMM workload, plain emitter, its own kernel name; the
  emitted launch requests ``OVER_LIMIT_LDS`` bytes of LDS."""
  from tensorforge_amd import workloads as W
  from tensorforge_amd.common.basic_types import FloatingPointType
  from tensorforge_amd.common.context import Context
  from tensorforge_amd.emit import launcher
  from tensorforge_amd.generators.generator import Generator
  w = W.gemm(FloatingPointType.DOUBLE)
  g = Generator(w.descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE), emitter='plain')
  g.set_kernel_name(PROBE_NAME)
  g.generate()
  em = dataclasses.replace(g.get_emitted(), lds_bytes=OVER_LIMIT_LDS)
  g._emitted = em
  g._launcher = launcher.launcher_source(g, em)
  return g


def so_path(directory=BUILD_DIR):
  return os.path.join(directory, 'liblaunch_error.so')


def build(directory=BUILD_DIR):
  from tensorforge_amd.emit.launcher import translation_unit
  from tensorforge_amd.runtime.build import compile_source
  g = generator()
  return compile_source(translation_unit([g]), 'launch_error', directory, keep_source=False)


if __name__ == '__main__':
  print(build(sys.argv[1] if len(sys.argv) > 1 else BUILD_DIR))
