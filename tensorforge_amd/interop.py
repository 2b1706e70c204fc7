"""Package-level entry points used by yateto (reference ``tensorforge/interop.py:1-23``)."""
import os

_HERE = os.path.dirname(os.path.realpath(__file__))


def get_cmake_path():
  """Directory of CMake package files. The package ships a ctypes/C-ABI library instead of a
  CMake package; the path names where ``tensorforge_aux.h`` and the C headers live."""
  return os.path.join(os.path.dirname(_HERE), 'include')


def print_cmake_path():
  print(get_cmake_path(), end='')


def get_routine_generator(yateto):
  from .frontend import yateto as fe
  return fe.YatetoFrontend


def get_version():
  with open(os.path.join(_HERE, 'VERSION')) as f:
    return f.read()


def print_version():
  print(get_version(), end='')


def use_fusedgemm_cost():
  return True
