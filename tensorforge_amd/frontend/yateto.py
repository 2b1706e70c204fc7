"""yateto adapter: turns yateto's tensor and operation records into descriptions, emits the
launcher call into yateto's generated C++ and registers the kernel with yateto's routine cache.

Restates the reference's frontend (``tensorforge/frontend/yateto.py:19-398``) against the same
duck-typed yateto objects (``op.name``, ``op.memoryLayout``, ``op.eqspp``, ``op.is_temporary``,
``op.values``, ``op.datatype``, ``arch.name/backend/typename/alignedLower/alignedUpper``).
"""
from ..common.aux import generate_tmp_tensor
from ..common.basic_types import Addressing, DataFlowDirection, FloatingPointType
from ..common.context import Context
from ..common.exceptions import GenerationError
from ..common.matrix.boundingbox import BoundingBox as BBox
from ..common.matrix.spp import FullSPP, ListSPP
from ..common.matrix.tensor import SubTensor, Tensor
from ..generators.descriptions import ElementwiseDescr, MultilinearDescr
from ..generators.generator import Generator as TensorForgeGenerator
from ..interface import YatetoInterface

_ADDRESSING = {'&': Addressing.NONE, 'n*N+o&': Addressing.STRIDED, 'n&+o&': Addressing.PTR_BASED,
               '': Addressing.SCALAR}


def _storage(layout):
  return layout.storage() if type(layout).__name__ == 'MemoryLayoutView' else layout


def _name_of(op):
  """yateto scalars and tensors expose ``name`` as a method, operation terms as an attribute."""
  return op.name() if callable(op.name) else op.name


class GpuKernelGeneratorV1:
  def __init__(self, arch):
    self._arch = arch
    self._cache = {}
    self._tmp_matrices = {}
    self._descr_list = []

  # ---- legacy gateway: yateto operation objects ----
  def add_operation(self, dest, ops, target, permute, add):
    self._cache_matrices(dest, ops, target, permute)
    aligned = self._can_be_aligned(dest, ops, target, permute)
    self._descr_list.append(MultilinearDescr(
      self.get_tensor(dest, aligned, list(range(len(dest.indices)))),
      [self.get_tensor(op, aligned, optarget) for op, optarget in zip(ops, target)],
      target, permute, add=add, strict_match=False, prefer_align=aligned))
    return 0

  # ---- pointwise statements ----
  def add_scalar(self, ops, statements, indices):
    """Pointwise statements (``optree``) over yateto tensors and scalars (reference
    ``frontend/yateto.py:127-149``): ``indices[i]`` holds the index ids of ``ops[i]``; every
    ``TensorVar``/``ScalarVar`` in ``statements`` is bound to its tensor by ``assignTensor``, a
    tensor at the bounding box of its nonzeros (``eqspp.nnzbounds()``)."""
    ids = {}
    for op, idx in zip(ops, indices):
      if _name_of(op) not in self._cache:      # records from add_tensor are already there
        self.make_tensor(op, False, None)
      ids[_name_of(op)] = idx

    def assigner(pre):
      name = _name_of(pre)
      if self.is_scalar(pre) and not hasattr(pre, 'eqspp'):
        if name not in self._cache:
          self.make_tensor(pre, False, None)
        ids[name] = []
        return SubTensor(self._cache[name], BBox([], [])), []
      eqspp = pre.eqspp() if callable(pre.eqspp) else pre.eqspp
      bounds = eqspp.nnzbounds()
      return (SubTensor(self._cache[name], BBox([lo for lo, _ in bounds], [hi + 1 for _, hi in bounds])),
              ids[name])

    for statement in statements:
      statement.assignTensor(assigner)
    self._descr_list.append(ElementwiseDescr(statements))
    return 0

  # ---- dictionary gateway ----
  def add_tensor(self, d):
    """Register a tensor record: storage kind 'full' / 'bbox' / 'spp', addressing string,
    values, flags (reference ``frontend/yateto.py:214-253``)."""
    name = d['name']
    datatype = FloatingPointType.ytt2enum(d['datatype'])
    shape = d['storage']['shape']
    kind = d['storage']['type']
    if d['addressing'] not in _ADDRESSING:
      raise ValueError(f'unknown addressing {d["addressing"]!r} of {name}')
    addressing = _ADDRESSING[d['addressing']]
    bbox, spp = None, None
    if kind == 'full':
      spp = FullSPP(shape)
    elif kind == 'bbox':
      starts, sizes = d['storage']['start'], d['storage']['sizes']
      bbox = BBox(list(starts), [s + n for s, n in zip(starts, sizes)])
      spp = FullSPP(shape)
    elif kind == 'spp':
      spp = ListSPP(d['storage']['entries'], shape)
    else:
      raise ValueError(f'unknown storage type {kind!r} of {name}')
    flags = d.get('flags', {})
    self._cache[name] = Tensor(shape, addressing, bbox, name, flags.get('temporary', False), spp,
                               d.get('values'), datatype)

  def tensor_ref(self, d):
    name = d['name']
    if name not in self._cache:
      raise KeyError(f'tensor {name} was not added')
    return SubTensor(self._cache[name], self._cache[name].bbox)

  def add_operation_new(self, d):
    if d['type'] != 'multilinear':
      raise NotImplementedError(f'operation type {d["type"]!r} is not generated for gfx950')
    result = self.tensor_ref(d['result'])
    args = [self.tensor_ref(a) for a in d['args']]
    add = d['linear']['add']
    self._descr_list.append(MultilinearDescr(result, args, d['target'], d['permute'], add=add))
    return 0

  # ---- yateto object helpers ----
  def is_scalar(self, op):
    return not hasattr(op, 'memoryLayout') and not isinstance(op, (float, int))

  def get_tensor(self, op, can_be_aligned, dims):
    if isinstance(op, (float, int)):
      return SubTensor(Tensor([], Addressing.SCALAR, data=[op]))
    if self.is_scalar(op):
      return SubTensor(self._cache[op.name()])
    tensor = self._cache[op.name]
    bounds = op.eqspp.nnzbounds()
    lower = [s for s, _ in bounds]
    upper = [e + 1 for _, e in bounds]
    layout = op.memoryLayout
    if type(layout).__name__ == 'MemoryLayoutView':
      rel = layout.relidx([0] * len(lower))
      lower = [x + rel[i] for i, x in enumerate(lower)]
      upper = [x + rel[i] for i, x in enumerate(upper)]
      layout = layout.storage()
    if can_be_aligned and lower and layout.alignedStride():
      lower[0] = self._arch.alignedLower(lower[0])
      upper[0] = min(self._arch.alignedUpper(upper[0]), layout.bbox()[0].stop)
    return SubTensor(tensor, BBox(lower, upper))

  def _can_be_aligned(self, dest, ops, target, permute):
    aligned = dest.memoryLayout.alignedStride()
    for i, _ in enumerate(ops):
      if 0 in target[i]:
        aligned = aligned and permute[i][0] == 0
    return aligned

  def make_tensor(self, op, can_be_aligned, dims):
    if isinstance(op, (float, int)):
      return Tensor([], Addressing.SCALAR, data=[op])
    if self.is_scalar(op):
      entry = Tensor([], Addressing.SCALAR, alias=op.name(), datatype=op.datatype)
      self._tmp_matrices[op.name()] = entry
      name = op.name()
    else:
      entry = self._matrix_from_yateto(op)
      name = op.name
    if not (name in self._cache and entry.is_same(self._cache[name])):
      self._cache[name] = entry

  def _matrix_from_yateto(self, tensor):
    layout = _storage(tensor.memoryLayout)
    if tensor.is_temporary:
      if tensor.name not in self._tmp_matrices:
        raise RuntimeError(f'expected tmp. tensor {tensor.name} to be cached '
                           'while code generation for fused-gemms')
      return self._tmp_matrices[tensor.name]
    addressing = tensor.addressing if getattr(tensor, 'addressing', None) is not None else \
      self._deduce_addressing(tensor)
    pattern = None
    if type(layout).__name__ != 'DenseMemoryLayout':
      rows = range(layout.bbox()[0].start, layout.bbox()[0].stop)
      cols = range(layout.bbox()[1].start, layout.bbox()[1].stop)
      pattern = layout.entries(rows, cols)
    return YatetoInterface.gen_matrix([rng.stop for rng in layout.bbox()], layout.bbox(),
                                      addressing=addressing, name=tensor.name,
                                      is_tmp=tensor.is_temporary, permute=None, pattern=pattern,
                                      values=tensor.values, datatype=tensor.datatype)

  @staticmethod
  def _deduce_addressing(term):
    if term.is_compute_constant:
      return Addressing.NONE
    if term.is_temporary:
      return Addressing.STRIDED
    return Addressing.PTR_BASED

  def _cache_matrices(self, dest, ops, target, permute):
    aligned = self._can_be_aligned(dest, ops, target, permute)
    if dest.is_temporary and dest in ops:
      raise ValueError('a temporary cannot be accumulated onto itself')
    for op, optarget in zip(ops, target):
      self.make_tensor(op, aligned, optarget)
    if dest.is_temporary:
      if dest.name not in self._tmp_matrices:
        tmp = generate_tmp_tensor([self.get_tensor(op, aligned, t) for op, t in zip(ops, target)],
                                  target, alias=dest.name)
        self._tmp_matrices[dest.name] = tmp
        self._cache[dest.name] = tmp
    else:
      self.make_tensor(dest, aligned, list(range(len(dest.indices))))

  # ---- code generation ----
  def generate(self, cpp, routine_cache):
    """One kernel for the whole list where it fuses (a contraction chain followed by pointwise
    statements on its destination, as the reference generates them); otherwise one kernel and
    one launcher call per run of descriptions of one kind, launched in order on the same stream.
    A chain temporary lives inside one kernel: using it across a split raises at generation."""
    fptype = FloatingPointType.str2enum(self._arch.typename) if hasattr(self._arch, 'typename') \
      else FloatingPointType.FLOAT
    segments = []
    for d in self._descr_list:
      pointwise = isinstance(d, ElementwiseDescr)
      if segments and segments[-1][0] == pointwise:
        segments[-1][1].append(d)
      else:
        segments.append((pointwise, [d]))
    if len(segments) == 2 and not segments[0][0]:
      context = Context(arch=self._arch.name, backend=self._arch.backend, fp_type=fptype)
      generator = TensorForgeGenerator(list(self._descr_list), context)
      generator.register()
      try:
        generator.generate()
      except GenerationError:
        generator = None
      if generator is not None:
        self._emit(cpp, routine_cache, generator, context)
        return
    for _, descrs in segments:
      context = Context(arch=self._arch.name, backend=self._arch.backend, fp_type=fptype)
      generator = TensorForgeGenerator(descrs, context)
      generator.register()
      generator.generate()
      self._emit(cpp, routine_cache, generator, context)

  def _emit(self, cpp, routine_cache, generator, context):
    cpp(self._gen_call_site(generator))
    routine_cache.addRoutine(generator.get_base_name(),
                             TensorForgeWriter(generator, context.get_vm().get_headers()))

  def _gen_call_site(self, generator):
    mat_name_map, offset_name_map = {}, {}
    for name, matrix in self._cache.items():
      if matrix.direction == DataFlowDirection.SOURCE and matrix.addressing != Addressing.SCALAR:
        datatype = FloatingPointType.str2enum(self._arch.typename) if matrix.datatype is None \
          else matrix.datatype
        mat_name_map[name] = f'const_cast<const {datatype}{matrix.addressing.to_pointer()}>({name})'
      else:
        mat_name_map[name] = name
      offset_name_map[name] = '0' if matrix.is_tmp or matrix.addressing == Addressing.NONE \
        else f'extraOffset_{name}'
    return generator.generate_call_site(mat_name_map, offset_name_map, 'numElements', 'flags',
                                        'streamPtr')


class TensorForgeWriter:
  """Appends kernel and launcher to yateto's routine file and returns the prototype
  (reference ``frontend/yateto.py:355-381``); the prototype is declared ``extern "C"``."""

  def __init__(self, generator, headers):
    self._headers = list(headers) + list(generator.get_helper_headers())
    self._generator = generator
    self._basename = generator.get_base_name()

  def target(self):
    return 'gpu'

  def __eq__(self, other):
    return isinstance(other, TensorForgeWriter) and self._basename == other._basename

  def __hash__(self):
    return hash(self._basename)

  def header(self, cpp):
    cpp.includes(self._headers)

  def __call__(self, routine_name, file_name):
    from ..emit.launcher import HEADER
    with open(file_name, 'a', encoding='utf-8') as f:
      f.write(HEADER)
      f.write(self._generator.get_kernel())
      f.write(self._generator.get_launcher())
    return 'extern "C" ' + self._generator.get_header()


class YatetoFrontend:
  def __init__(self, arch):
    self.generator = GpuKernelGeneratorV1(arch)

  def generate(self, cpp, cache):
    self.generator.generate(cpp, cache)

  def add_linear_operation(self, dest, ops, target, permute, add):
    return self.generator.add_operation(dest, ops, target, permute, add)

  def add_operation(self, description):
    return self.generator.add_operation_new(description)

  def add_tensor(self, tensor):
    return self.generator.add_tensor(tensor)

  def add_scalar(self, ops, statements, indices):
    return self.generator.add_scalar(ops, statements, indices)
