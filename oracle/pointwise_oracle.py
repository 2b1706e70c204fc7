"""NumPy oracle for ``ElementwiseDescr`` lists (pointwise statement trees).

TEST INFRASTRUCTURE ONLY: the checker that pointwise kernels are compared against; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may use it. It imports
nothing from ``tensorforge_amd``: statement and expression nodes are told apart by class name.

Formulation (independent of the kernel's three-phase form): the statements run one after the
other, each over all index points of all active elements at once, reading from and writing to
the launcher-layout arrays themselves. For independent points (what the reference's per-lane loop
already assumes) this is the reference's per-point order.

Semantics restated from the reference:

* index ranges: every index id ``-d-1`` ranges over ``[0, min(bbox extents it indexes))``
  (``tensorforge/generators/optree.py:238-247``, ``backend/instructions/compute/elementwise.py:31-41``);
* entry read or written: ``n + bbox.lower()`` in tensor coordinates, addressed
  ``sum((idx - tensor.bbox.lower) * colmajor_stride)`` from the element's base
  (``optree.py:251-262``, ``backend/symbol.py:304-312``); element base STRIDED
  ``e * prod(shape) + extraOffset``, PTR_BASED ``ptr[e] + extraOffset``, NONE shared
  (``backend/instructions/ptr_manip.py:19-65``);
* statements in order per point; ``IfNode`` runs its body where the condition holds, ``WhileNode``
  repeats it while the condition holds (``optree.py:488-548``);
* elements with ``flags[e] == 0`` and elements ``>= numElements`` are untouched
  (``generators/generator.py:494-505``).

Operations follow C++ (the reference emits C++ expressions, ``common/vm/lexic/cuda_lexic.py:113-205``):
integer ``/`` truncates, ``%`` keeps the dividend's sign, math functions work in floating point,
comparisons give bool, ``&``/``|``/``^`` are logical on two bools and bitwise on integers,
``round`` rounds half away from zero, float to integer conversion truncates. Floating values are
carried in FP64 whatever the kernel's type.

Pinned by ``tests/test_pointwise.py``: hand-computed values (C++ integer division and
remainder, the while loop, ternaries, bounding boxes), Python's ``math`` functions at chosen
points, the closed form of the update and the return map's invariants (mean stress kept, yielded
nodes on the yield surface, idempotence).
"""
import math

import numpy as np

from .chain_oracle import OracleInputs   # noqa: F401  (same input container)

_INT_NAMES = ('BOOL', 'INT', 'UINT', 'LONG', 'ULONG')


def _strides(shape):
  out, s = [], 1
  for x in shape:
    out.append(s)
    s *= x
  return out


def _kind_of_type(fptype):
  """'bool', 'int' or 'float' for a FloatingPointType (None: the kernel's floating type)."""
  if fptype is None:
    return 'float'
  name = fptype.name
  if name == 'BOOL':
    return 'bool'
  if name in _INT_NAMES:
    return 'int'
  return 'float'


class _Value:
  """An array over (active element, point) with its C++ kind."""

  def __init__(self, data, kind):
    self.data = data
    self.kind = kind


def _num(v):
  """Bool promotes to int in arithmetic."""
  if v.kind == 'bool':
    return _Value(v.data.astype(np.int64), 'int')
  return v


def _arith_kind(a, b):
  return 'float' if 'float' in (a.kind, b.kind) else 'int'


def _as(v, kind):
  if kind == 'float':
    return v.data.astype(np.float64)
  if kind == 'int':
    if v.kind == 'float':
      return np.trunc(v.data).astype(np.int64)
    return v.data.astype(np.int64)
  return v.data != 0


class _Run:
  def __init__(self, descrs, inp):
    self.inp = inp
    self.program = [s for d in descrs for s in d.oplist]
    n = inp.num_elements
    elems = np.arange(n)
    if inp.flags is not None:
      elems = elems[np.asarray(inp.flags[:n]) != 0]
    self.elems = elems
    ranges = {}
    for var in self._tensor_vars():
      for dim, idx in enumerate(var.indices):
        size = var.tensor.bbox.upper()[dim] - var.tensor.bbox.lower()[dim]
        ranges[idx] = min(ranges.get(idx, size), size)
    rank = len(ranges)
    self.extents = [ranges[-d - 1] for d in range(rank)]
    # index grid: point p (first index fastest) -> n_d
    pts = int(np.prod(self.extents)) if rank else 1
    grid = np.indices(list(reversed(self.extents))) if rank else np.zeros((0, 1), dtype=np.int64)
    self.n = [grid[rank - 1 - d].reshape(-1) for d in range(rank)] if rank else []
    self.points = pts
    self.shape = (len(elems), pts)
    self.temps = {}
    # PTR_BASED lists become 2-D arrays while running
    self.ptr_based = {}
    for t, v in inp.arrays.items():
      if isinstance(v, list):
        self.ptr_based[t] = np.stack([np.asarray(x, dtype=np.float64) for x in v])

  def _tensor_vars(self):
    out = []

    def nodes(node):
      yield node
      for c in getattr(node, 'operands', []) or []:
        yield from nodes(c)

    def walk(stmts):
      for s in stmts:
        name = type(s).__name__
        if name == 'Assignment':
          for root in (s.dest, s.optree):
            out.extend(x for x in nodes(root) if type(x).__name__ == 'TensorVar')
        else:
          out.extend(x for x in nodes(s.condition) if type(x).__name__ == 'TensorVar')
          walk(s.subassignments)
    walk(self.program)
    return out

  # ---- addressing ----
  def _flat(self, var):
    """Flat offsets (active element, point) of a TensorVar inside its element."""
    t = var.tensor.tensor
    strides = _strides(t.shape)
    off = np.zeros(self.points, dtype=np.int64)
    for dim, idx in enumerate(var.indices):
      d = -idx - 1
      absolute = self.n[d] + var.tensor.bbox.lower()[dim]
      off += (absolute - t.bbox.lower()[dim]) * strides[dim]
    return off

  def _data_table(self, t):
    vol = int(np.prod(t.shape)) if len(t.shape) else 1
    table = np.zeros(vol)
    strides = _strides(t.shape)
    for idx in np.ndindex(*t.shape):
      v = t.value(idx)
      if v is not None:
        table[sum((i - o) * s for i, o, s in zip(idx, t.bbox.lower(), strides))] = v
    return table

  def read(self, var):
    t = var.tensor.tensor
    off = self._flat(var)
    kind = _kind_of_type(t.datatype)
    if t.data is not None and t not in self.inp.arrays:
      return _Value(np.broadcast_to(self._data_table(t)[off], self.shape).copy(), kind)
    extra = self.inp.offsets.get(t, 0)
    addressing = str(t.addressing)
    if addressing == 'pointer_based':
      arr = self.ptr_based[t]
      data = arr[self.elems[:, None], extra + off[None, :]]
    elif addressing == 'none':
      data = np.broadcast_to(np.asarray(self.inp.arrays[t])[extra + off], self.shape).copy()
    else:
      vol = int(np.prod(t.shape))
      data = np.asarray(self.inp.arrays[t])[self.elems[:, None] * vol + extra + off[None, :]]
    return _Value(np.asarray(data), kind)

  def write(self, var, value, mask):
    t = var.tensor.tensor
    off = self._flat(var)
    kind = _kind_of_type(t.datatype)
    data = _as(value, kind).astype(np.float64)
    extra = self.inp.offsets.get(t, 0)
    addressing = str(t.addressing)
    e_idx = np.broadcast_to(self.elems[:, None], self.shape)[mask]
    p_off = np.broadcast_to(off[None, :], self.shape)[mask]
    vals = data[mask]
    if addressing == 'pointer_based':
      self.ptr_based[t][e_idx, extra + p_off] = vals
    elif addressing == 'none':
      self.inp.arrays[t][extra + p_off] = vals
    else:
      vol = int(np.prod(t.shape))
      self.inp.arrays[t][e_idx * vol + extra + p_off] = vals

  # ---- expressions ----
  def scalar(self, var):
    t = var.tensor.tensor
    kind = _kind_of_type(t.datatype)
    if t.data is not None:
      v = t.data[0] if not isinstance(t.data, dict) else next(iter(t.data.values()))
    else:
      v = self.inp.arrays[t]
    return _Value(np.full(self.shape, v, dtype=np.float64 if kind == 'float' else np.int64), kind)

  def expr(self, node):
    name = type(node).__name__
    if name == 'TensorVar':
      return self.read(node)
    if name == 'ScalarVar':
      return self.scalar(node)
    if name == 'TempVar':
      if id(node) not in self.temps:
        kind = _kind_of_type(node.datatype)
        self.temps[id(node)] = _Value(np.zeros(self.shape), kind)
      v = self.temps[id(node)]
      return _Value(v.data.copy(), v.kind)
    if name == 'Immediate':
      fp = node.fptype.name
      kind = 'bool' if fp == 'BOOL' else ('int' if fp in _INT_NAMES else 'float')
      return _Value(np.full(self.shape, node.value,
                            dtype={'bool': bool, 'int': np.int64, 'float': np.float64}[kind]), kind)
    ops = [self.expr(o) for o in node.operands]
    if name == 'CastOpNode':
      kind = _kind_of_type(node.targetType)
      return _Value(_as(ops[0], kind), kind)
    if name == 'ConditionalOpNode':
      c, a, b = ops
      kind = 'bool' if a.kind == b.kind == 'bool' else _arith_kind(_num(a), _num(b))
      return _Value(np.where(_as(c, 'bool'), _as(a, kind), _as(b, kind)), kind)
    return self.operation(node.optype.name, ops)

  def operation(self, op, v):
    if op == 'COPY':
      return v[0]
    if op in ('ADD', 'SUB', 'MUL', 'DIV', 'MOD'):
      a, b = _num(v[0]), _num(v[1])
      kind = _arith_kind(a, b)
      x, y = _as(a, kind), _as(b, kind)
      if op == 'ADD':
        return _Value(x + y, kind)
      if op == 'SUB':
        return _Value(x - y, kind)
      if op == 'MUL':
        return _Value(x * y, kind)
      if kind == 'int':
        q = np.sign(x) * np.sign(y) * (np.abs(x) // np.abs(y))     # C++: toward zero
        return _Value(q if op == 'DIV' else x - q * y, 'int')
      return _Value(x / y if op == 'DIV' else np.fmod(x, y), 'float')
    if op == 'NEG':
      a = _num(v[0])
      return _Value(-a.data, a.kind)
    if op == 'RCP':
      a = _num(v[0])
      if a.kind == 'int':
        return self.operation('DIV', [_Value(np.ones(self.shape, dtype=np.int64), 'int'), a])
      return _Value(1.0 / a.data, 'float')
    cmp = {'LT': np.less, 'LE': np.less_equal, 'GT': np.greater, 'GE': np.greater_equal,
           'EQ': np.equal, 'NEQ': np.not_equal}
    if op in cmp:
      a, b = _num(v[0]), _num(v[1])
      kind = _arith_kind(a, b)
      return _Value(cmp[op](_as(a, kind), _as(b, kind)), 'bool')
    if op in ('AND', 'OR', 'XOR'):
      if v[0].kind == 'bool' and v[1].kind == 'bool':
        f = {'AND': np.logical_and, 'OR': np.logical_or, 'XOR': np.not_equal}[op]
        return _Value(f(v[0].data, v[1].data), 'bool')
      f = {'AND': np.bitwise_and, 'OR': np.bitwise_or, 'XOR': np.bitwise_xor}[op]
      return _Value(f(_as(v[0], 'int'), _as(v[1], 'int')), 'int')
    if op == 'NOT':
      if v[0].kind == 'bool':
        return _Value(~v[0].data.astype(bool), 'bool')
      return _Value(np.invert(_as(v[0], 'int')), 'int')
    x = [_as(a, 'float') for a in v]
    with np.errstate(all='ignore'):
      if op == 'MIN':
        return _Value(np.fmin(x[0], x[1]), 'float')
      if op == 'MAX':
        return _Value(np.fmax(x[0], x[1]), 'float')
      if op == 'POW':
        return _Value(np.power(x[0], x[1]), 'float')
      if op == 'ROUND':
        t = np.trunc(x[0])
        return _Value(np.where(np.abs(x[0] - t) >= 0.5, t + np.sign(x[0]), t), 'float')
      if op == 'SIGN':
        return _Value(np.sign(x[0]), 'float')
      if op == 'RSQRT':
        return _Value(1.0 / np.sqrt(x[0]), 'float')
      if op == 'RCBRT':
        return _Value(1.0 / np.cbrt(x[0]), 'float')
      if op == 'GAMMA':
        return _Value(np.vectorize(math.gamma)(x[0]), 'float')
      if op == 'ERF':
        return _Value(np.vectorize(math.erf)(x[0]), 'float')
      unary = {'SQRT': np.sqrt, 'CBRT': np.cbrt, 'ABS': np.abs, 'CEIL': np.ceil,
               'FLOOR': np.floor, 'TRUNC': np.trunc, 'LOGP1': np.log1p, 'EXPM1': np.expm1,
               'LOG': np.log, 'EXP': np.exp, 'SIN': np.sin, 'COS': np.cos, 'TAN': np.tan,
               'ASIN': np.arcsin, 'ACOS': np.arccos, 'ATAN': np.arctan, 'SINH': np.sinh,
               'COSH': np.cosh, 'TANH': np.tanh, 'ASINH': np.arcsinh, 'ACOSH': np.arccosh,
               'ATANH': np.arctanh}
      if op in unary:
        return _Value(unary[op](x[0]), 'float')
    raise NotImplementedError(op)

  # ---- statements ----
  def execute(self, stmts, mask, max_iterations=100000):
    for s in stmts:
      name = type(s).__name__
      if name == 'Assignment':
        value = self.expr(s.optree)
        dname = type(s.dest).__name__
        if dname == 'TensorVar':
          self.write(s.dest, value, mask)
        elif dname == 'TempVar':
          cur = self.expr(s.dest)
          kind = cur.kind
          self.temps[id(s.dest)] = _Value(np.where(mask, _as(value, kind), cur.data), kind)
        else:
          raise ValueError(f'cannot assign to {dname}')
      elif name == 'IfNode':
        cond = _as(self.expr(s.condition), 'bool')
        self.execute(s.subassignments, mask & cond)
      elif name == 'WhileNode':
        for _ in range(max_iterations):
          live = mask & _as(self.expr(s.condition), 'bool')
          if not live.any():
            break
          self.execute(s.subassignments, live)
        else:
          raise RuntimeError('while loop did not end')
      else:
        raise ValueError(f'unknown statement {name}')

  def run(self):
    if self.points and len(self.elems):
      self.execute(self.program, np.ones(self.shape, dtype=bool))
    for t, arr in self.ptr_based.items():
      self.inp.arrays[t] = [row.copy() for row in arr]
    return self.inp.arrays


def run(descrs, inp):
  """Execute the ``ElementwiseDescr`` list ``descrs`` on ``inp`` (arrays modified in place)."""
  return _Run(descrs, inp).run()
