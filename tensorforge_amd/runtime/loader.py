"""ctypes access to ``libtensorforge_rt.so`` and to generated launchers.

The launcher's argument types come from the generator's prototype
(``Generator.launcher_signature``), so a call has exactly the reference ABI: device pointers,
``unsigned`` extra offsets, ``size_t numElements``, ``unsigned* flags``, ``void* streamPtr``.
Loading fails loudly when an object is missing: there is no CPU fallback.
"""
import ctypes
import os

from .build import runtime_path

_RT = None


def runtime():
  """The runtime library, loaded once with ``RTLD_GLOBAL`` so kernel objects resolve it."""
  global _RT
  if _RT is None:
    path = runtime_path()
    if not os.path.exists(path):
      raise RuntimeError(f'{path} is missing: build the package first '
                         '(python -c "import __graft_entry__ as g; g.build()")')
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    lib.tensorforge_last_error.restype = ctypes.c_int
    lib.tensorforge_last_error_where.restype = ctypes.c_char_p
    lib.tensorforge_error_string.restype = ctypes.c_char_p
    lib.tensorforge_error_string.argtypes = [ctypes.c_int]
    lib.tensorforge_record_error.argtypes = [ctypes.c_int, ctypes.c_char_p]
    lib.tensorforge_synch_device.restype = ctypes.c_int
    lib.tensorforge_synch_device.argtypes = [ctypes.c_void_p]
    lib.tensorforge_copy.restype = ctypes.c_int
    lib.tensorforge_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_void_p]
    for fn in ('tensorforge_graph_begin', 'tensorforge_graph_launch', 'tensorforge_graph_destroy'):
      getattr(lib, fn).restype = ctypes.c_int
    lib.tensorforge_graph_begin.argtypes = [ctypes.c_void_p]
    lib.tensorforge_graph_end.restype = ctypes.c_int
    lib.tensorforge_graph_end.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.tensorforge_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.tensorforge_graph_destroy.argtypes = [ctypes.c_void_p]
    _RT = lib
  return _RT


_CTYPES = {
  'double': ctypes.c_double, 'float': ctypes.c_float, 'unsigned': ctypes.c_uint,
  'size_t': ctypes.c_size_t, 'int': ctypes.c_int, 'long': ctypes.c_long,
  'bool': ctypes.c_bool, 'unsigned int': ctypes.c_uint, 'unsigned long': ctypes.c_ulong,
}


def _ctype(c_type: str):
  if c_type.endswith('*'):
    return ctypes.c_void_p
  return _CTYPES[c_type]


class Kernel:
  """A generated launcher, callable with the reference's argument list."""

  def __init__(self, generator, lib):
    self.generator = generator
    self.name = f'launcher_{generator.get_base_name()}'
    try:
      self.fn = getattr(lib, self.name)
    except AttributeError as err:
      raise RuntimeError(f'{self.name} is not exported by {lib._name}') from err
    self.signature = generator.launcher_signature()
    self.fn.argtypes = [_ctype(t) for t, _ in self.signature]
    self.fn.restype = None
    self.lib = lib

  def __call__(self, *args):
    if len(args) != len(self.signature):
      raise TypeError(f'{self.name} takes {len(self.signature)} arguments '
                      f'({", ".join(n for _, n in self.signature)}), given {len(args)}')
    conv = []
    for (t, _), a in zip(self.signature, args):
      if t.endswith('*'):
        conv.append(None if a is None or a == 0 else ctypes.c_void_p(int(a)))
      else:
        conv.append(a)
    self.fn(*conv)
    err = runtime().tensorforge_last_error()
    if err != 0:
      where = runtime().tensorforge_last_error_where().decode()
      msg = runtime().tensorforge_error_string(err).decode()
      runtime().tensorforge_clear_error()
      raise RuntimeError(f'{where}: HIP error {err} ({msg})')


def load_kernels(so_path, generators):
  """Load ``so_path`` and bind one :class:`Kernel` per generator."""
  if not os.path.exists(so_path):
    raise RuntimeError(f'{so_path} is missing: the kernels were not built')
  runtime()
  lib = ctypes.CDLL(so_path)
  return [Kernel(g, lib) for g in generators]
