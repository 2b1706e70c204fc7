"""Enumerations shared by the description API and the gfx950 emitter.

Same names, values and string forms as the reference
(``tensorforge/common/basic_types.py:4-169``): the string forms feed the kernel-name hash, so they
must not change.
"""
import enum


class DataFlowDirection(enum.Enum):
  SOURCE = 0
  SINK = 1
  SOURCESINK = 2


_ADDR_TO_STR = None
_STR_TO_ADDR = None


class Addressing(enum.Enum):
  """How a batched operand is found for element ``e``.

  * ``NONE``: one array shared by all elements (``T*``);
  * ``STRIDED``: element ``e`` at ``ptr + e * volume + extraOffset`` (``T*``);
  * ``PTR_BASED``: element ``e`` at ``ptr[e] + extraOffset`` (``T**``);
  * ``SCALAR``: a by-value kernel argument.
  """
  NONE = 0
  STRIDED = 1
  PTR_BASED = 2
  SCALAR = 3

  @classmethod
  def addr2ptr_type(cls, addr_type):
    return {cls.NONE: '*', cls.STRIDED: '*', cls.PTR_BASED: '**', cls.SCALAR: ''}[addr_type]

  @classmethod
  def addr2str(cls, addr):
    return {cls.NONE: 'none', cls.STRIDED: 'strided', cls.PTR_BASED: 'pointer_based',
            cls.SCALAR: 'scalar'}[addr]

  @classmethod
  def str2addr(cls, string):
    table = {'none': cls.NONE, 'strided': cls.STRIDED, 'pointer_based': cls.PTR_BASED,
             'scalar': cls.SCALAR}
    if string not in table:
      raise ValueError('arg must be either none, strided, pointer_based, or scalar; '
                       f'given: {string}')
    return table[string]

  def to_pointer(self):
    return self.addr2ptr_type(self)

  def __str__(self):
    return self.addr2str(self)


class StridedAddressing:
  """Strided addressing with an explicit element stride and offset (reference
  ``basic_types.py:50-59``)."""

  def __init__(self, offset, stride=None):
    self.offset = offset
    self.stride = stride

  def to_pointer(self):
    return Addressing.STRIDED.to_pointer()


class FloatingPointType(enum.Enum):
  FLOAT = 0
  DOUBLE = 1
  HALF = 2
  BFLOAT16 = 3
  BOOL = 10
  INT = 20
  UINT = 21
  LONG = 22
  ULONG = 23

  def size(self):
    return {FloatingPointType.FLOAT: 4, FloatingPointType.DOUBLE: 8, FloatingPointType.HALF: 2,
            FloatingPointType.BFLOAT16: 2, FloatingPointType.BOOL: 1, FloatingPointType.INT: 4,
            FloatingPointType.UINT: 4, FloatingPointType.LONG: 8,
            FloatingPointType.ULONG: 8}[self]

  def literal(self, value):
    if self == FloatingPointType.FLOAT:
      return f'{float(value):.16}f'
    if self == FloatingPointType.DOUBLE:
      return f'{float(value):.16}'
    if self == FloatingPointType.HALF:
      return f'static_cast<_Float16>({float(value):.16})'
    if self == FloatingPointType.BFLOAT16:
      return f'static_cast<__bf16>({float(value):.16})'
    if self == FloatingPointType.BOOL:
      return 'true' if value else 'false'
    if self == FloatingPointType.INT:
      return f'{int(value)}'
    if self == FloatingPointType.UINT:
      return f'{int(value)}u'
    if self == FloatingPointType.LONG:
      return f'{int(value)}L'
    return f'{int(value)}UL'

  @classmethod
  def as_str(cls, fp):
    return {cls.FLOAT: 'float', cls.DOUBLE: 'double', cls.HALF: 'half', cls.BFLOAT16: 'bfloat16',
            cls.BOOL: 'bool', cls.INT: 'int', cls.LONG: 'long', cls.UINT: 'unsigned int',
            cls.ULONG: 'unsigned long'}[fp]

  @classmethod
  def str2enum(cls, as_str: str):
    return {'float': cls.FLOAT, 'double': cls.DOUBLE, 'half': cls.HALF, 'bfloat16': cls.BFLOAT16,
            'bool': cls.BOOL, 'int': cls.INT, 'long': cls.LONG, 'unsigned int': cls.UINT,
            'unsigned long': cls.ULONG}[as_str]

  @classmethod
  def ytt2enum(cls, as_str: str):
    return {'f32': cls.FLOAT, 'f64': cls.DOUBLE, 'f16': cls.HALF, 'bf16': cls.BFLOAT16,
            'bool': cls.BOOL, 'i8': cls.INT, 'i16': cls.INT, 'i32': cls.INT, 'i64': cls.LONG,
            'u32': cls.UINT, 'u64': cls.ULONG}[as_str]

  def __str__(self):
    return self.as_str(self)


class GeneralLexicon:
  """Fixed identifiers used in generated launchers (reference ``basic_types.py:157-169``)."""
  BATCH_ID_NAME = 'batchId'
  NUM_ELEMENTS = 'numElements'
  EXTRA_OFFSET = '_extraOffset'
  STREAM_PTR_STR = 'streamPtr'
  ALPHA_SYMBOL_NAME = 'alpha'
  BETA_SYMBOL_NAME = 'beta'
  FLAGS_NAME = 'flags'
  GLOBAL_MEM_PREFIX = 'glb_'
  TOTAL_SHR_MEM = 'totalShrMem'
  LOCAL_SHR_MEM = 'localShrMem'
  SHR_MEM_REGION_PREFIX = 'shrRegion'
  REG_NAME = 'reg'
