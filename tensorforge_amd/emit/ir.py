"""Lowering of a description list into a per-element chain program.

Every ``MultilinearDescr`` becomes one :class:`Product`: index ranges, operand placement
(kernel argument, chain temporary, by-value scalar, baked constant) and sparsity resolved at
generation time. Both gfx950 emitters read only this program.

Index ranges follow the reference's analysis (``backend/instructions/compute/multilinear.py:58-96``):
an output or contracted index ranges over the intersection of the bounding boxes of every operand
dimension mapped to it, in absolute tensor coordinates. Chain temporaries are stored relative to
the lower corner of the range that produced them.
"""
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from ..common.basic_types import Addressing, DataFlowDirection, FloatingPointType
from ..common.exceptions import GenerationError
from ..common.matrix.tensor import Tensor

SRC_GLOBAL = 'global'    # a launcher argument (pointer)
SRC_TMP = 'tmp'          # a chain temporary, never in HBM
SRC_SCALAR = 'scalar'    # a by-value launcher argument
SRC_DATA = 'data'        # values baked into the kernel


@dataclass
class GlobalArg:
  """One global-scope tensor: what the launcher takes for it."""
  tensor: Tensor
  name: str
  kind: str                    # SRC_GLOBAL / SRC_SCALAR / SRC_DATA
  addressing: Addressing
  dtype: FloatingPointType
  const: bool

  @property
  def has_offset(self):
    return self.kind != SRC_SCALAR and self.addressing not in (Addressing.NONE, Addressing.SCALAR)

  @property
  def per_element(self):
    return self.kind == SRC_GLOBAL and self.addressing in (Addressing.STRIDED, Addressing.PTR_BASED)

  @property
  def element_stride(self):
    """Distance between consecutive elements for STRIDED: the full dense volume, also for
    packed sparse tensors (reference ``backend/instructions/ptr_manip.py:39-45``)."""
    return self.tensor.get_real_volume()


@dataclass
class Factor:
  """One operand of a product, with each storage dim bound to an index label."""
  tensor: Tensor
  src: str
  labels: List[Tuple[str, int]]          # per storage dim: ('n', t) or ('k', c)
  arg: Optional[GlobalArg] = None
  origin: Tuple[int, ...] = ()           # index stored at offset 0 (tensor bbox lower / tmp lo)
  strides: Tuple[int, ...] = ()          # column-major strides of the stored array
  packed: Optional[Dict[Tuple[int, ...], int]] = None   # sparse: index -> packed position
  values: Optional[Dict[Tuple[int, ...], float]] = None  # baked constants

  @property
  def rank(self):
    return len(self.labels)

  def is_sparse(self):
    return self.packed is not None

  def offset(self, index):
    """Storage offset of ``index`` (absolute coordinates) or None for a structural zero."""
    if self.packed is not None:
      return self.packed.get(tuple(index))
    return sum((i - o) * s for i, o, s in zip(index, self.origin, self.strides))

  def structural_nonzero(self, index):
    if self.values is not None:
      v = self.values.get(tuple(index))
      return v is not None and v != 0.0
    if self.packed is not None:
      return tuple(index) in self.packed
    return True


@dataclass
class Product:
  dest: Tensor
  dest_src: str
  dest_arg: Optional[GlobalArg]
  dest_origin: Tuple[int, ...]
  dest_strides: Tuple[int, ...]
  n_ranges: List[Tuple[int, int]]
  k_ranges: List[Tuple[int, int]]
  factors: List[Factor]
  scalars: List[Factor]
  add: bool
  beta: float
  descr: object = None

  @property
  def out_rank(self):
    return len(self.n_ranges)

  def out_shape(self):
    return [hi - lo for lo, hi in self.n_ranges]

  def is_empty(self):
    """No output entry or no contraction term: nothing to multiply."""
    return any(hi <= lo for lo, hi in self.n_ranges + self.k_ranges)

  def out_empty(self):
    """No output entry: the product writes nothing. (An empty contraction with outputs still
    writes them: ``alpha * 0 + beta * dest``, zero when not accumulating.)"""
    return any(hi <= lo for lo, hi in self.n_ranges)

  def dest_offset(self, n):
    return sum((i - o) * s for i, o, s in zip(n, self.dest_origin, self.dest_strides))


@dataclass
class ChainProgram:
  args: List[GlobalArg]
  products: List[Product]
  fp_type: FloatingPointType
  tmp_shapes: Dict[int, Tuple[int, ...]] = field(default_factory=dict)
  # pointwise statements applied to the chain's destination before it is stored
  epilogue: object = None

  def arg_by_tensor(self, tensor):
    for a in self.args:
      if a.tensor is tensor:
        return a
    return None

  def flops_per_element(self):
    """Nonzero multiply-adds x2, plus one add per accumulated output entry."""
    total = 0
    for p in self.products:
      total += 2 * count_terms(p)
      if p.add:
        total += p.out_shape()[0] * _prod(p.out_shape()[1:])
    return total


def _prod(xs):
  r = 1
  for x in xs:
    r *= x
  return r


def colmajor_strides(shape):
  strides, s = [], 1
  for extent in shape:
    strides.append(s)
    s *= extent
  return tuple(strides)


def iterate_box(ranges):
  """All index tuples of a box, first dim fastest."""
  if not ranges:
    yield ()
    return
  import itertools
  axes = [range(lo, hi) for lo, hi in ranges]
  for rev in itertools.product(*reversed(axes)):
    yield tuple(reversed(rev))


def count_terms(p: Product):
  """Number of (n, k) terms of ``p`` whose factors are all structurally nonzero."""
  if p.is_empty():
    return 0
  count = 0
  for n in iterate_box(p.n_ranges):
    for k in iterate_box(p.k_ranges):
      if all(f.structural_nonzero(factor_index(f, n, k)) for f in p.factors):
        count += 1
  return count


def factor_index(f: Factor, n, k):
  return tuple(n[i] if kind == 'n' else k[i] for kind, i in f.labels)


def global_scope_kind(tensor: Tensor):
  """Symbol kind of a global-scope tensor (reference ``generators/generator.py:330-347``)."""
  if tensor.has_values() and len(tensor.get_values()) < 16:
    return SRC_DATA
  if tensor.addressing == Addressing.SCALAR:
    return SRC_SCALAR
  return SRC_GLOBAL


def build_args(global_tensors, fp_type):
  args = []
  for t in global_tensors:
    kind = global_scope_kind(t)
    args.append(GlobalArg(tensor=t, name=t.name, kind=kind, addressing=t.addressing,
                          dtype=t.datatype if t.datatype is not None else fp_type,
                          const=t.direction == DataFlowDirection.SOURCE))
  return args


def lower(descr_list, global_tensors, fp_type) -> ChainProgram:
  """Lower ``descr_list`` (already named by the generator) into a :class:`ChainProgram`."""
  from ..generators.descriptions import MultilinearDescr
  args = build_args(global_tensors, fp_type)
  by_id = {id(a.tensor): a for a in args}
  tmp_origin: Dict[int, Tuple[int, ...]] = {}
  tmp_shape: Dict[int, Tuple[int, ...]] = {}
  products = []
  for descr in descr_list:
    if not isinstance(descr, MultilinearDescr):
      raise GenerationError(f'the gfx950 emitter generates multilinear descriptions only, '
                            f'given {type(descr).__name__}')
    products.append(_lower_one(descr, by_id, tmp_origin, tmp_shape))
  return ChainProgram(args=args, products=products, fp_type=fp_type, tmp_shapes=tmp_shape)


def _lower_one(descr, by_id, tmp_origin, tmp_shape) -> Product:
  n_rank = 0
  for t in descr.target:
    for x in t:
      n_rank = max(n_rank, x + 1)
  k_count = 0
  for t in descr.target:
    for x in t:
      if x < 0:
        k_count = max(k_count, -x)
  n_ranges = [None] * n_rank
  k_ranges = [None] * k_count

  def narrow(ranges, i, lo, hi):
    if ranges[i] is None:
      ranges[i] = (lo, hi)
    else:
      ranges[i] = (max(ranges[i][0], lo), min(ranges[i][1], hi))

  factors, scalars = [], []
  for i, op in enumerate(descr.ops):
    tensor = op.tensor
    target = descr.target[i]
    rank = len(tensor.shape)
    if len(target) != op.bbox.rank():
      raise GenerationError(f'operand {tensor.name}: target {target} does not match its rank '
                            f'{op.bbox.rank()}')
    perm = descr.logical_permute(i)
    labels = [None] * rank
    for j, t in enumerate(target):
      p = perm[j]
      lo, hi = op.bbox.lower()[p], op.bbox.upper()[p]
      if t >= 0:
        labels[p] = ('n', t)
        narrow(n_ranges, t, lo, hi)
      else:
        labels[p] = ('k', -t - 1)
        narrow(k_ranges, -t - 1, lo, hi)
    factor = _make_factor(tensor, labels, by_id, tmp_origin, tmp_shape)
    if factor.src in (SRC_SCALAR, SRC_DATA) and all(s == 1 for s in tensor.shape):
      scalars.append(factor)
    else:
      factors.append(factor)

  for t, r in enumerate(n_ranges):
    if r is None:
      raise GenerationError(f'output dimension {t} of {descr.dest.tensor.name} is not produced '
                            f'by any operand')
  for c, r in enumerate(k_ranges):
    if r is None:
      raise GenerationError(f'contracted index {-c - 1} is missing from every operand')
  # disjoint boxes give an empty range: keep its lower end, size zero
  k_ranges = [(lo, max(lo, hi)) for lo, hi in k_ranges]

  dest = descr.dest.tensor
  if len(dest.shape) != n_rank and not dest.is_tmp:
    raise GenerationError(f'destination {dest.name} has rank {len(dest.shape)}, the product '
                          f'has rank {n_rank}')
  if dest.is_tmp:
    if id(dest) in tmp_origin:
      # accumulate into (or overwrite) an existing temporary: keep its coordinates
      origin = tmp_origin[id(dest)]
      shape = tmp_shape[id(dest)]
      n_ranges = [(max(lo, o), max(max(lo, o), min(hi, o + s)))
                  for (lo, hi), o, s in zip(n_ranges, origin, shape)]
    else:
      if descr.add:
        raise GenerationError(f'cannot accumulate into temporary {dest.name} before it is produced')
      origin = tuple(lo for lo, _ in n_ranges)
      shape = tuple(max(hi - lo, 0) for lo, hi in n_ranges)
      tmp_origin[id(dest)] = origin
      tmp_shape[id(dest)] = shape
    return Product(dest=dest, dest_src=SRC_TMP, dest_arg=None, dest_origin=origin,
                   dest_strides=colmajor_strides(tmp_shape[id(dest)]), n_ranges=n_ranges,
                   k_ranges=k_ranges, factors=factors, scalars=scalars, add=descr.add,
                   beta=descr.beta, descr=descr)
  arg = by_id.get(id(dest))
  if arg is None or arg.kind != SRC_GLOBAL:
    raise GenerationError(f'destination {dest.name} must be a pointer argument')
  if dest.is_sparse():
    raise GenerationError(f'destination {dest.name} is sparse; sparse results are not supported')
  view = descr.dest.bbox
  n_ranges = [(max(lo, vl, tl), min(hi, vu, tu)) for (lo, hi), vl, vu, tl, tu in
              zip(n_ranges, view.lower(), view.upper(), dest.bbox.lower(), dest.bbox.upper())]
  n_ranges = [(lo, max(lo, hi)) for lo, hi in n_ranges]
  return Product(dest=dest, dest_src=SRC_GLOBAL, dest_arg=arg, dest_origin=tuple(dest.bbox.lower()),
                 dest_strides=colmajor_strides(dest.shape), n_ranges=n_ranges, k_ranges=k_ranges,
                 factors=factors, scalars=scalars, add=descr.add, beta=descr.beta, descr=descr)


def _make_factor(tensor, labels, by_id, tmp_origin, tmp_shape) -> Factor:
  if tensor.is_tmp:
    if id(tensor) not in tmp_origin:
      raise GenerationError(f'temporary {tensor.name} is used before it is produced')
    shape = tmp_shape[id(tensor)]
    return Factor(tensor=tensor, src=SRC_TMP, labels=labels, origin=tuple([0] * len(shape)),
                  strides=colmajor_strides(shape))
  arg = by_id[id(tensor)]
  if arg.kind == SRC_DATA:
    values = {}
    for idx in iterate_box([(0, s) for s in tensor.shape]):
      v = tensor.value(idx)
      if v is not None:
        values[idx] = float(v)
    return Factor(tensor=tensor, src=SRC_DATA, labels=labels, arg=arg, values=values)
  if arg.kind == SRC_SCALAR:
    return Factor(tensor=tensor, src=SRC_SCALAR, labels=labels, arg=arg)
  packed = None
  if tensor.is_sparse():
    packed = {}
    for idx in tensor.spp.nonzeros():
      packed[tuple(idx)] = tensor.linear_index(idx)
  return Factor(tensor=tensor, src=SRC_GLOBAL, labels=labels, arg=arg,
                origin=tuple(tensor.bbox.lower()), strides=colmajor_strides(tensor.shape),
                packed=packed)
