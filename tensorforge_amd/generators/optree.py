"""Pointwise statement trees: the body of an ``ElementwiseDescr``.

Same classes, constructors and helper functions as the reference
(``tensorforge/generators/optree.py:10-785``): ``Assignment(dest, optree)``, ``TensorVar``,
``ScalarVar``, ``TempVar``, ``Immediate``, ``LexicOpNode``, ``ConditionalOpNode``, ``CastOpNode``,
``IfNode``, ``WhileNode``; ``assign``, ``conditional``, ``ternary``, ``loop``, ``tensor``,
``scalar``, ``imm``, ``temp``, ``cast`` and the math/compare/bit functions, with the reference's
rewrites of trivial products, quotients and powers (``optree.py:633-720``). Python operators on
nodes build trees (``a * b + 1``, ``x < y``, ``-x``, ``abs(x)``).

Index convention (reference ``optree.py:238-258``, ``compute/elementwise.py:31-57``): a
``TensorVar`` carries one index id per storage dimension; ids are ``-1, -2, ...`` and id ``-d-1``
is loop variable ``n{d}``. Every id ranges over ``[0, min(bbox extents of the dimensions it
indexes))``, and the entry read or written is ``n + bbox.lower()`` in the tensor's own
coordinates. Index ids come from the yateto frontend's ``assignTensor`` (``TensorVar(None, None,
pretensor)``) or directly: ``tensor(sub_tensor, indices=[-1, -2])``.

What differs from the reference, each deliberate:

* Every node has a deterministic ``str()`` (the reference's ``Assignment`` has none, so the
  kernel-name hash of an ``ElementwiseDescr`` includes object addresses and changes per run).
* A node is evaluated where it is used: a node object used twice, or re-evaluated by a
  ``WhileNode``, is computed again (the reference caches the first ``const auto`` and so never
  re-evaluates a loop condition, ``optree.py:435-442,537-546``).
* Code generation, symbol assignment and declaration live in the gfx950 emitter
  (:mod:`tensorforge_amd.emit.pointwise`), not on the nodes.
"""
import builtins as _builtins
import math as _math
from typing import List, Union

from ..common.basic_types import Addressing, FloatingPointType
from ..common.exceptions import GenerationError
from ..common.matrix.tensor import SubTensor, Tensor
from ..common.operation import Operation


class Node:
  """An expression. Operators build trees; ``==`` builds a comparison, so nodes hash by
  identity."""
  __hash__ = object.__hash__

  def tensors(self, intensors=True, outtensors=True):
    return []

  def pretensors(self, intensors=True, outtensors=True):
    return []

  def getRanges(self, ranges):
    return ranges

  def assignTensor(self, assigner):
    pass

  def children(self):
    return []

  def __add__(self, other):
    return add(self, other)

  def __radd__(self, other):
    return add(other, self)

  def __sub__(self, other):
    return sub(self, other)

  def __rsub__(self, other):
    return sub(other, self)

  def __mul__(self, other):
    return mul(self, other)

  def __rmul__(self, other):
    return mul(other, self)

  def __truediv__(self, other):
    return div(self, other)

  def __rtruediv__(self, other):
    return div(other, self)

  def __mod__(self, other):
    return mod(self, other)

  def __rmod__(self, other):
    return mod(other, self)

  def __or__(self, other):
    return bitor(self, other)

  def __ror__(self, other):
    return bitor(other, self)

  def __and__(self, other):
    return bitand(self, other)

  def __rand__(self, other):
    return bitand(other, self)

  def __xor__(self, other):
    return bitxor(self, other)

  def __rxor__(self, other):
    return bitxor(other, self)

  def __pow__(self, other):
    return pow(self, other)

  def __rpow__(self, other):
    return pow(other, self)

  def __lt__(self, other):
    return complt(self, other)

  def __le__(self, other):
    return comple(self, other)

  def __gt__(self, other):
    return compgt(self, other)

  def __ge__(self, other):
    return compge(self, other)

  def __eq__(self, other):
    return compeq(self, other)

  def __ne__(self, other):
    return compne(self, other)

  def __neg__(self):
    return neg(self)

  def __pos__(self):
    return self

  def __abs__(self):
    return abs(self)

  def __invert__(self):
    return bitnot(self)


class Variable(Node):
  """A node that can also be assigned to."""


class Statement:
  __hash__ = object.__hash__

  def tensors(self, intensors=True, outtensors=True):
    return []

  def pretensors(self, intensors=True, outtensors=True):
    return []

  def getRanges(self, ranges):
    return ranges

  def assignTensor(self, assigner):
    pass


class Assignment(Statement):
  def __init__(self, dest: Variable, optree: Node):
    if not isinstance(dest, (TensorVar, ScalarVar, TempVar)):
      raise GenerationError(f'cannot assign to {type(dest).__name__}')
    self.dest = dest
    self.optree = optree

  def tensors(self, intensors=True, outtensors=True):
    out = []
    if intensors:
      out += self.optree.tensors(intensors, outtensors)
    if outtensors:
      out += self.dest.tensors(intensors, outtensors)
    return out

  def pretensors(self, intensors=True, outtensors=True):
    out = []
    if intensors:
      out += self.optree.pretensors(intensors, outtensors)
    if outtensors:
      out += self.dest.pretensors(intensors, outtensors)
    return out

  def getRanges(self, ranges):
    ranges = self.dest.getRanges(ranges)
    return self.optree.getRanges(ranges)

  def assignTensor(self, assigner):
    self.dest.assignTensor(assigner)
    self.optree.assignTensor(assigner)

  def __str__(self):
    return f'{self.dest} = {self.optree}'


def _as_subtensor(x):
  if isinstance(x, Tensor):
    return SubTensor(x)
  return x


class TensorVar(Variable):
  """One entry of a batched tensor per index point. ``indices`` (an extension of the reference
  constructor) binds the index ids directly instead of through ``assignTensor``."""

  def __init__(self, tensor, slicing, pretensor=None, indices=None):
    self.tensor = _as_subtensor(tensor)
    self.slicing = slicing
    self.pretensor = pretensor
    self.indices = None if indices is None else [int(i) for i in indices]
    self.offset = None if self.tensor is None else list(self.tensor.bbox.lower())
    if self.tensor is not None and self.indices is not None:
      self._check()

  def _check(self):
    rank = len(self.tensor.bbox.sizes())
    if len(self.indices) != rank:
      raise GenerationError(f'{len(self.indices)} index ids for a rank-{rank} tensor '
                            f'{self.tensor}')
    for i in self.indices:
      if i >= 0:
        raise GenerationError(f'index ids are negative (-1 is the first loop index), got {i}')

  def tensors(self, intensors=True, outtensors=True):
    return [self.tensor]

  def pretensors(self, intensors=True, outtensors=True):
    return [self.pretensor]

  def getRanges(self, ranges):
    for dim, idx in enumerate(self.indices):
      size = self.tensor.bbox.size(dim)
      if idx not in ranges:
        ranges[idx] = (0, size)
      ranges[idx] = (0, _builtins.min(ranges[idx][1], size))
    return ranges

  def assignTensor(self, assigner):
    self.tensor, self.indices = assigner(self.pretensor)
    self.tensor = _as_subtensor(self.tensor)
    self.indices = [int(i) for i in self.indices]
    self.offset = list(self.tensor.bbox.lower())
    self._check()

  def __str__(self):
    return f'{self.tensor}{self.indices}'


class ScalarVar(Variable):
  """A by-value scalar argument (``Addressing.SCALAR``), or a constant scalar (``data``)."""

  def __init__(self, tensor, slicing, pretensor=None):
    self.tensor = _as_subtensor(tensor)
    self.slicing = slicing
    self.pretensor = pretensor
    self.indices = [] if tensor is not None else None

  def tensors(self, intensors=True, outtensors=True):
    return [self.tensor]

  def assignTensor(self, assigner):
    self.tensor, self.indices = assigner(self.pretensor)
    self.tensor = _as_subtensor(self.tensor)

  def __str__(self):
    return f'{self.tensor}'


class TempVar(Variable):
  """A per-point local variable. ``name`` is given by the ``ElementwiseDescr`` that holds it
  (first-appearance order) unless set here."""

  def __init__(self, datatype=None, name=None):
    self.datatype = datatype
    self.name = name

  def __str__(self):
    dt = '' if self.datatype is None else f':{self.datatype}'
    return f'{self.name or "tmp"}{dt}'


class Immediate(Variable):
  def __init__(self, value, fptype: FloatingPointType):
    self.value = value
    self.fptype = fptype

  def __str__(self):
    if self.fptype == FloatingPointType.BOOL:
      return 'true' if self.value else 'false'
    if self.fptype in (FloatingPointType.FLOAT, FloatingPointType.DOUBLE):
      return repr(float(self.value))
    return str(int(self.value))


class OpNode(Node):
  def __init__(self, operands: List[Node]):
    self.operands = operands

  def children(self):
    return list(self.operands)

  def getRanges(self, ranges):
    for op in self.operands:
      ranges = op.getRanges(ranges)
    return ranges

  def tensors(self, intensors=True, outtensors=True):
    return [t for op in self.operands for t in op.tensors()]

  def pretensors(self, intensors=True, outtensors=True):
    return [t for op in self.operands for t in op.pretensors()]

  def assignTensor(self, assigner):
    for op in self.operands:
      op.assignTensor(assigner)


class LexicOpNode(OpNode):
  def __init__(self, operands: List[Node], optype: Operation):
    super().__init__(operands)
    self.optype = optype

  def __str__(self):
    return f'{self.optype}({", ".join(str(o) for o in self.operands)})'


class ConditionalOpNode(OpNode):
  def __str__(self):
    return f'select({", ".join(str(o) for o in self.operands)})'


class CastOpNode(OpNode):
  def __init__(self, operands: List[Node], targetType: FloatingPointType):
    super().__init__(operands)
    self.targetType = targetType

  def __str__(self):
    return f'cast<{self.targetType}>({self.operands[0]})'


class _Block(Statement):
  def __init__(self, condition: Node, subassignments: List[Statement]):
    self.condition = condition
    self.subassignments = list(subassignments)

  def getRanges(self, ranges):
    ranges = self.condition.getRanges(ranges)
    for s in self.subassignments:
      ranges = s.getRanges(ranges)
    return ranges

  def assignTensor(self, assigner):
    self.condition.assignTensor(assigner)
    for s in self.subassignments:
      s.assignTensor(assigner)

  def tensors(self, intensors=True, outtensors=True):
    out = self.condition.tensors(intensors, outtensors) if intensors else []
    return out + [t for s in self.subassignments for t in s.tensors(intensors, outtensors)]

  def pretensors(self, intensors=True, outtensors=True):
    out = self.condition.pretensors(intensors, outtensors) if intensors else []
    return out + [t for s in self.subassignments for t in s.pretensors(intensors, outtensors)]

  def _body(self):
    return '; '.join(str(s) for s in self.subassignments)


class IfNode(_Block):
  def __str__(self):
    return f'if ({self.condition}) {{ {self._body()} }}'


class WhileNode(_Block):
  def __str__(self):
    return f'while ({self.condition}) {{ {self._body()} }}'


def walk_statements(statements):
  """Every statement, depth first, in program order."""
  for s in statements:
    yield s
    if isinstance(s, _Block):
      yield from walk_statements(s.subassignments)


def walk_nodes(node):
  yield node
  for c in node.children():
    yield from walk_nodes(c)


def statement_nodes(s):
  """Top-level expression trees of one statement (not those of nested statements)."""
  if isinstance(s, Assignment):
    return [s.dest, s.optree]
  if isinstance(s, _Block):
    return [s.condition]
  return []


def scalarblock(statements: List[Statement]):
  raise GenerationError('scalar regions belong to the yateto frontend: pass the statements to '
                        'YatetoFrontend.add_scalar')


BaseType = Union[Node, float, int, bool]


def imm(value, fptype):
  return Immediate(value, fptype)


def tensor(x, slicing=None, indices=None):
  """A tensor read/write: ``x`` is a yateto node (bound later by ``assignTensor``) or a
  ``Tensor``/``SubTensor`` with its ``indices``."""
  if isinstance(x, (Tensor, SubTensor)):
    if indices is None:
      raise GenerationError('tensor(): a Tensor or SubTensor needs its index ids')
    return TensorVar(x, slicing, indices=indices)
  return TensorVar(None, slicing, x)


def scalar(x):
  if isinstance(x, (Tensor, SubTensor)):
    t = x.tensor if isinstance(x, SubTensor) else x
    if t.addressing != Addressing.SCALAR:
      raise GenerationError(f'scalar(): {t} is not a SCALAR-addressed tensor')
    return ScalarVar(x, None)
  return ScalarVar(None, None, x)


def immc(x: BaseType):
  if isinstance(x, bool):
    return imm(x, FloatingPointType.BOOL)
  if isinstance(x, float):
    return imm(x, FloatingPointType.FLOAT)
  if isinstance(x, int):
    return imm(x, FloatingPointType.INT)
  if isinstance(x, Node):
    return x
  raise GenerationError(f'cannot use {type(x).__name__} in a pointwise expression')


def assign(target, source: BaseType):
  if not isinstance(target, Node):
    target = tensor(target)
  return Assignment(target, immc(source))


def conditional(condition: BaseType, subnodes: List[Statement]):
  return IfNode(immc(condition), subnodes)


def ternary(condition: BaseType, yesnode: BaseType, nonode: BaseType):
  return ConditionalOpNode([immc(condition), immc(yesnode), immc(nonode)])


def loop(condition: BaseType, subnodes: List[Statement]):
  return WhileNode(immc(condition), subnodes)


def temp(datatype=None):
  return TempVar(datatype)


def cast(x: BaseType, fptype: FloatingPointType):
  return CastOpNode([immc(x)], fptype)


def _unary(op):
  return lambda x: LexicOpNode([immc(x)], op)


def _binary(op):
  return lambda x, y: LexicOpNode([immc(x), immc(y)], op)


cos = _unary(Operation.COS)
sin = _unary(Operation.SIN)
tan = _unary(Operation.TAN)
acos = _unary(Operation.ACOS)
asin = _unary(Operation.ASIN)
atan = _unary(Operation.ATAN)
cosh = _unary(Operation.COSH)
sinh = _unary(Operation.SINH)
tanh = _unary(Operation.TANH)
acosh = _unary(Operation.ACOSH)
asinh = _unary(Operation.ASINH)
atanh = _unary(Operation.ATANH)
sqrt = _unary(Operation.SQRT)
cbrt = _unary(Operation.CBRT)
exp = _unary(Operation.EXP)
log = _unary(Operation.LOG)
expm1 = _unary(Operation.EXPM1)
logp1 = _unary(Operation.LOGP1)
round = _unary(Operation.ROUND)   # noqa: A001 - reference name
# operations the reference's enum names but gives no builder for
floor = _unary(Operation.FLOOR)
ceil = _unary(Operation.CEIL)
trunc = _unary(Operation.TRUNC)
sign = _unary(Operation.SIGN)
erf = _unary(Operation.ERF)
gamma = _unary(Operation.GAMMA)
rsqrt = _unary(Operation.RSQRT)
rcbrt = _unary(Operation.RCBRT)
rcp = _unary(Operation.RCP)
neg = _unary(Operation.NEG)
abs = _unary(Operation.ABS)       # noqa: A001 - reference name
bitnot = _unary(Operation.NOT)
max = _binary(Operation.MAX)      # noqa: A001 - reference name
min = _binary(Operation.MIN)      # noqa: A001 - reference name
add = _binary(Operation.ADD)
sub = _binary(Operation.SUB)
mod = _binary(Operation.MOD)
bitand = _binary(Operation.AND)
bitor = _binary(Operation.OR)
bitxor = _binary(Operation.XOR)
complt = _binary(Operation.LT)
comple = _binary(Operation.LE)
compgt = _binary(Operation.GT)
compge = _binary(Operation.GE)
compeq = _binary(Operation.EQ)
compne = _binary(Operation.NEQ)


def _is_value(node, *values):
  return isinstance(node, Immediate) and node.fptype != FloatingPointType.BOOL and \
    any(node.value == v for v in values)


def mul(x: BaseType, y: BaseType):
  """``1·y -> y``, ``-1·y -> -y`` and mirrored (reference ``optree.py:633-648``)."""
  xc, yc = immc(x), immc(y)
  if _is_value(xc, 1):
    return yc
  if _is_value(xc, -1):
    return neg(yc)
  if _is_value(yc, 1):
    return xc
  if _is_value(yc, -1):
    return neg(xc)
  return LexicOpNode([xc, yc], Operation.MUL)


def div(x: BaseType, y: BaseType):
  """``1/y -> rcp(y)``, ``-1/y -> rcp(-y)``, ``x/1 -> x``, ``x/-1 -> -x`` (reference
  ``optree.py:650-665``)."""
  xc, yc = immc(x), immc(y)
  if _is_value(xc, 1):
    return LexicOpNode([yc], Operation.RCP)
  if _is_value(xc, -1):
    return LexicOpNode([neg(yc)], Operation.RCP)
  if _is_value(yc, 1):
    return xc
  if _is_value(yc, -1):
    return neg(xc)
  return LexicOpNode([xc, yc], Operation.DIV)


def pow(x: BaseType, y: BaseType):  # noqa: A001 - reference name
  """Constant exponents become cheaper forms (reference ``optree.py:681-704``): 2 -> ``x·x``,
  0.5 -> sqrt, -0.5 -> rsqrt, 1/3 -> cbrt, -1/3 -> rcbrt, -1 -> rcp, 1 -> x; base e -> exp,
  base 1 -> 1."""
  xc, yc = immc(x), immc(y)
  if isinstance(yc, Immediate) and yc.fptype != FloatingPointType.BOOL:
    v = yc.value
    if v == 2:
      return LexicOpNode([xc, xc], Operation.MUL)
    if v == 0.5:
      return LexicOpNode([xc], Operation.SQRT)
    if v == -0.5:
      return LexicOpNode([xc], Operation.RSQRT)
    if v == 1 / 3:
      return LexicOpNode([xc], Operation.CBRT)
    if v == -1 / 3:
      return LexicOpNode([xc], Operation.RCBRT)
    if v == -1:
      return LexicOpNode([xc], Operation.RCP)
    if v == 1:
      return xc
  if isinstance(xc, Immediate) and xc.fptype != FloatingPointType.BOOL:
    if xc.value == _math.e:
      return LexicOpNode([yc], Operation.EXP)
    if xc.value == 1:
      return xc
  return LexicOpNode([xc, yc], Operation.POW)


def matmul(x: BaseType, y: BaseType):
  raise GenerationError('matmul inside a pointwise tree: use a MultilinearDescr')


def einsum(op: str, *args):
  raise GenerationError('einsum inside a pointwise tree: use a MultilinearDescr')
