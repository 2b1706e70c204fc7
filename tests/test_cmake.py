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


@pytest.mark.skipif(shutil.which('cmake') is None, reason='cmake is not installed')
def test_reference_integration_project(tmp_path):
  """The reference's own CMake integration check (tests/cmake_integration/CMakeLists.txt:1-11):
  CMAKE_PREFIX_PATH from print_cmake_path(), find_package(Tensorforge REQUIRED), the include
  directories, a C++ file including tensorforge_aux.h."""
  (tmp_path / 'CMakeLists.txt').write_text(
    'cmake_minimum_required(VERSION 3.5)\nproject(cmake_integration)\n'
    'execute_process(COMMAND python3 -c "import tensorforge; tensorforge.print_cmake_path()"\n'
    '                OUTPUT_VARIABLE TENSORFORGE_PATH)\n'
    'set(CMAKE_PREFIX_PATH "${TENSORFORGE_PATH}" ${CMAKE_MODULE_PATH})\n'
    'find_package(Tensorforge REQUIRED)\n'
    'add_executable(${CMAKE_PROJECT_NAME} ./main.cpp)\n'
    'target_include_directories(${CMAKE_PROJECT_NAME} PRIVATE ${TensorForge_INCLUDE_DIRS})\n')
  (tmp_path / 'main.cpp').write_text(
    '#include <tensorforge_aux.h>\n#include <iostream>\n'
    'int main() { std::cout << "hello tensorforge" << std::endl; return 0; }\n')
  env = dict(os.environ, PYTHONPATH=ROOT)
  for cmd in (['cmake', '-S', str(tmp_path), '-B', str(tmp_path / 'b')],
              ['cmake', '--build', str(tmp_path / 'b')]):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
  r = subprocess.run([str(tmp_path / 'b' / 'cmake_integration')], capture_output=True, text=True)
  assert r.returncode == 0 and 'hello tensorforge' in r.stdout
