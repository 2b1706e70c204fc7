"""hipcc driver for gfx950 shared objects.

Everything is compiled in-tree into ``tensorforge_amd/lib`` so the objects travel with the
package. Device code is built with packed-FP32 instructions disabled
(``-target-feature -packed-fp32-ops``): the kernels compute in plain FP32/FP64 VALU or MFMA.
"""
import hashlib
import os
import shutil
import subprocess
import tempfile

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, 'lib')
CSRC_DIR = os.path.join(PKG_DIR, 'csrc')
REPO_DIR = os.path.dirname(PKG_DIR)
INCLUDE_DIR = os.path.join(REPO_DIR, 'include')
RUNTIME_NAME = 'libtensorforge_rt.so'


def hipcc():
  path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
  if not os.path.exists(path):
    raise RuntimeError('hipcc not found; the gfx950 kernels need ROCm to build')
  return path


def hipcc_flags():
  return ['-O3', '--offload-arch=gfx950', '-std=c++17', '-fPIC', '-shared',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wno-unused-command-line-argument']


def runtime_path():
  return os.path.join(LIB_DIR, RUNTIME_NAME)


def _run(cmd, what):
  proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
  if proc.returncode != 0:
    raise RuntimeError(f'{what} failed ({proc.returncode}):\n{" ".join(cmd)}\n{proc.stdout[-6000:]}')
  return proc.stdout


def _atomic_compile(cmd_without_out, out_path, what):
  os.makedirs(os.path.dirname(out_path), exist_ok=True)
  fd, tmp = tempfile.mkstemp(prefix=os.path.basename(out_path) + '.tmp', dir=os.path.dirname(out_path))
  os.close(fd)
  try:
    _run(cmd_without_out + ['-o', tmp], what)
    os.replace(tmp, out_path)
  finally:
    if os.path.exists(tmp):
      os.remove(tmp)


RUNTIME_SOURCES = ('tensorforge_rt.hip', 'tensorforge_mlp.hip')


def build_runtime():
  srcs = [os.path.join(CSRC_DIR, s) for s in RUNTIME_SOURCES]
  _atomic_compile([hipcc()] + hipcc_flags() + [f'-I{INCLUDE_DIR}'] + srcs, runtime_path(),
                  'runtime build')
  return runtime_path()


def source_digest(source: str) -> str:
  return hashlib.sha1(source.encode()).hexdigest()[:12]


def compile_source(source: str, name: str, out_dir: str = LIB_DIR, keep_source=True,
                   extra_flags=()) -> str:
  """Compile one HIP translation unit into ``<out_dir>/lib<name>.so`` linked against the
  runtime library (found next to it at load time through ``$ORIGIN``)."""
  os.makedirs(out_dir, exist_ok=True)
  src_path = os.path.join(out_dir, f'{name}.hip')
  with open(src_path, 'w') as f:
    f.write(source)
  out = os.path.join(out_dir, f'lib{name}.so')
  rt_dir = os.path.dirname(runtime_path())
  cmd = [hipcc()] + hipcc_flags() + list(extra_flags) + [
    f'-I{INCLUDE_DIR}', src_path, f'-L{rt_dir}', '-ltensorforge_rt', f'-Wl,-rpath,{rt_dir}',
    '-Wl,-rpath,$ORIGIN']
  _atomic_compile(cmd, out, f'kernel build of {name}')
  if not keep_source:
    os.remove(src_path)
  return out


def compile_generators(generators, name: str, out_dir: str = LIB_DIR, extra_flags=()) -> str:
  """Generate (if needed) and compile ``generators`` into one shared object (``extra_flags``:
  further hipcc options, e.g. ``--offload-compress`` for objects kept in bulk)."""
  from ..emit.launcher import translation_unit
  for g in generators:
    if g.get_kernel() is None:
      g.generate()
  if not os.path.exists(runtime_path()):
    build_runtime()
  return compile_source(translation_unit(generators), name, out_dir, extra_flags=extra_flags)


def host_cxx():
  path = shutil.which('g++')
  if path is None:
    raise RuntimeError('g++ not found; the omptarget launchers need a host C++ compiler')
  return path


def compile_host_source(source: str, name: str, out_dir: str) -> str:
  """Compile a host C++/OpenMP translation unit (the omptarget backend's launchers) into
  ``<out_dir>/lib<name>.so``: no fused multiply-add, so each output's sum is rounded term by
  term as the reference's OpenMP kernels round it."""
  os.makedirs(out_dir, exist_ok=True)
  src_path = os.path.join(out_dir, f'{name}.cpp')
  with open(src_path, 'w') as f:
    f.write(source)
  out = os.path.join(out_dir, f'lib{name}.so')
  _atomic_compile([host_cxx(), '-O2', '-fopenmp', '-ffp-contract=off', '-std=c++17', '-fPIC',
                   '-shared', src_path], out, f'host build of {name}')
  return out
