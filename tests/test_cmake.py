"""The CMake package (find_package(TensorForge)) and the generated routine-file format, with a
C++ consumer built out of tree. CPU only: configure, generate and compile for gfx950."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'tests', 'cmake_consumer'))


def test_cmake_path_holds_the_package():
  import tensorforge
  path = tensorforge.get_cmake_path()
  assert os.path.isfile(os.path.join(path, 'TensorForgeConfig.cmake'))
  assert os.path.isfile(os.path.join(path, 'TensorForgeConfigVersion.cmake'))


def test_routine_file_and_call_sites(tmp_path):
  """The chain and its pointwise statements as one fused kernel appended to the routine file,
  its extern "C" prototype and call site; the file compiles for gfx950 on its own."""
  import build_consumer
  calls = build_consumer.generate(str(tmp_path))
  assert len(calls) == 1
  assert 'const_cast<const double*>(A)' in calls[0] and calls[0].endswith(
    'dt, numElements, flags, streamPtr);')
  text = (tmp_path / 'kernels.h').read_text()
  assert text.count('extern "C" void launcher_kernel_') == 1
  src = tmp_path / 'kernels.cpp'
  from tensorforge_amd.runtime.build import INCLUDE_DIR, hipcc, hipcc_flags
  flags = [f for f in hipcc_flags() if f not in ('-shared', '-fPIC')]
  r = subprocess.run([hipcc()] + flags + [f'-I{INCLUDE_DIR}', '-x', 'hip', '-c', str(src), '-o',
                                          str(tmp_path / 'kernels.o')],
                     capture_output=True, text=True)
  assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.skipif(shutil.which('cmake') is None, reason='cmake is not installed')
def test_consumer_configures_and_builds(tmp_path):
  import build_consumer
  exe = build_consumer.build(str(tmp_path / 'build'))
  assert os.access(exe, os.X_OK)
  flags = (tmp_path / 'build' / 'CMakeFiles' / 'consumer.dir' / 'flags.make').read_text()
  assert '--offload-arch=gfx950' in flags and 'packed-fp32-ops' in flags


@pytest.mark.skipif(shutil.which('cmake') is None, reason='cmake is not installed')
def test_other_backends_are_refused(tmp_path):
  (tmp_path / 'CMakeLists.txt').write_text(
    'cmake_minimum_required(VERSION 3.24)\nproject(x LANGUAGES CXX)\n'
    'find_package(TensorForge REQUIRED)\n')
  import tensorforge
  r = subprocess.run(['cmake', '-S', str(tmp_path), '-B', str(tmp_path / 'b'),
                      f'-DTensorForge_DIR={tensorforge.get_cmake_path()}',
                      '-DDEVICE_BACKEND=cuda'], capture_output=True, text=True)
  assert r.returncode != 0 and 'Allowed: hip' in (r.stdout + r.stderr)


RT_SYMBOLS = ('_ZN11tensorforge8checkErrERKNSt7__cxx1112basic_stringIcSt11char_traitsIcESaIcEEEi',
              '_ZN11tensorforge11synchDeviceEPv')


def test_runtime_exports_the_reference_shim():
  """libtensorforge_rt.so exports tensorforge::checkErr(const std::string&, int) and
  tensorforge::synchDevice(void*) out of line, with the signatures of the reference's
  tensorforge_aux.h, next to the C status functions."""
  from tensorforge_amd.runtime.build import runtime_path
  if not os.path.exists(runtime_path()):
    pytest.skip('runtime library not built')
  r = subprocess.run(['nm', '-D', '--defined-only', runtime_path()], capture_output=True, text=True)
  assert r.returncode == 0
  for sym in RT_SYMBOLS + ('tensorforge_last_error', 'tensorforge_synch_device'):
    assert sym in r.stdout, sym


@pytest.mark.skipif(shutil.which('cmake') is None, reason='cmake is not installed')
def test_prefix_path_integration(tmp_path):
  """A project that finds the package the way the reference's integration check does
  (CMAKE_PREFIX_PATH from ``print_cmake_path()``, then ``find_package(Tensorforge)`` in that
  spelling) and calls the shim through ``tensorforge_aux.h``: it must configure, compile and
  link ``CHECK_ERR`` and ``tensorforge::synchDevice`` against the runtime library."""
  from tensorforge_amd.runtime.build import runtime_path
  if not os.path.exists(runtime_path()):
    pytest.skip('runtime library not built')
  (tmp_path / 'CMakeLists.txt').write_text(
    'cmake_minimum_required(VERSION 3.16)\n'
    'project(tf_prefix_check CXX)\n'
    f'execute_process(COMMAND "{sys.executable}" -c "import tensorforge as t; t.print_cmake_path()"\n'
    '                OUTPUT_VARIABLE TF_PREFIX OUTPUT_STRIP_TRAILING_WHITESPACE)\n'
    'list(APPEND CMAKE_PREFIX_PATH "${TF_PREFIX}")\n'
    'find_package(Tensorforge REQUIRED)\n'
    'add_executable(shim_check shim_check.cpp)\n'
    'target_include_directories(shim_check PRIVATE ${TensorForge_INCLUDE_DIRS} /opt/rocm/include)\n'
    'target_compile_definitions(shim_check PRIVATE __HIP_PLATFORM_AMD__)\n'
    'target_link_libraries(shim_check PRIVATE TensorForge::TensorForge)\n')
  (tmp_path / 'shim_check.cpp').write_text(
    '// links the reference shim\'s declarations against libtensorforge_rt.so\n'
    '#include <tensorforge_aux.h>\n'
    '#include <cstdio>\n'
    'int main(int argc, char**) {\n'
    '  if (argc > 1) { tensorforge::synchDevice(); CHECK_ERR; }\n'
    '  std::printf("%d\\n", tensorforge_last_error());\n'
    '  return 0;\n'
    '}\n')
  env = dict(os.environ, PYTHONPATH=ROOT)
  for cmd in (['cmake', '-S', str(tmp_path), '-B', str(tmp_path / 'b')],
              ['cmake', '--build', str(tmp_path / 'b')]):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
  exe = tmp_path / 'b' / 'shim_check'
  assert os.access(exe, os.X_OK)
  r = subprocess.run(['nm', '-u', str(exe)], capture_output=True, text=True)
  for sym in RT_SYMBOLS:
    assert sym in r.stdout, sym
