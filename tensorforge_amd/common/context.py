"""``Context`` and ``Options``: the generation settings a description list is compiled under.

Same constructor and methods as the reference (``tensorforge/common/context.py:6-54``). Only
``arch='gfx950', backend='hip'`` is accepted; other targets raise ``ValueError``.
"""
from math import ceil

from .basic_types import FloatingPointType
from .hw import hw_descr_factory


class Options:
  def __init__(self, exact_contraction_length=False, align_shr_mem=True,
               enable_sync_block_opt=True):
    self.exact_contraction_length: bool = exact_contraction_length
    self.align_shr_mem: bool = align_shr_mem
    self.enable_sync_block_opt = enable_sync_block_opt


class _VM:
  """Holds the hardware description; exists so ``context.get_vm().get_hw_descr()`` keeps
  working for callers written against the reference (``common/vm/vm.py:7-42``)."""

  def __init__(self, hw, fp_type):
    self._hw = hw
    self._fp = fp_type

  def get_hw_descr(self):
    return self._hw

  def fp_as_str(self):
    return FloatingPointType.as_str(self._fp)

  def get_headers(self):
    return ['tensorforge_aux.h', 'hip/hip_runtime.h']


class Context:
  def __init__(self, arch: str, backend: str, fp_type: FloatingPointType,
               options: Options = None):
    if fp_type not in (FloatingPointType.FLOAT, FloatingPointType.DOUBLE):
      raise RuntimeError(f'unknown fp_type. Allowed float, double, given {fp_type}')
    self._hw = hw_descr_factory(arch, backend)
    self.arch = arch
    self.backend = backend
    self.fp_type = fp_type
    self._options = options if options is not None else Options()
    self._vm = _VM(self._hw, fp_type)

  def set_fp_type(self, fp_type: FloatingPointType):
    self.fp_type = fp_type
    self._vm = _VM(self._hw, fp_type)

  def fp_as_str(self):
    return FloatingPointType.as_str(self.fp_type)

  def get_vm(self):
    return self._vm

  def get_hw(self):
    return self._hw

  def get_user_options(self):
    return self._options

  def align(self, num):
    """Round ``num`` up to a multiple of one wavefront's worth of words of ``fp_type``
    (32 doubles or 64 floats on gfx950), as the reference does (``context.py:38-44``)."""
    align_length = (self._hw.vec_unit_length * self._hw.hw_fp_word_size) / self.fp_type.size()
    return int(ceil(num / align_length) * align_length)

  def align_range(self, begin, end):
    assert end > begin
    factor = self._hw.mem_access_align_size / self.fp_type.size()
    aligned_begin = begin - begin % factor
    aligned_end = end + (factor - end % factor) % factor
    return int(aligned_begin), int(aligned_end)
