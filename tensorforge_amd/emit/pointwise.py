"""gfx950 emitter for ``ElementwiseDescr``: pointwise statements over every index point of every
element of the batch.

Reference behaviour (``backend/instructions/compute/elementwise.py:11-70``,
``generators/optree.py:166-785``): one 64-lane group per element, the first index spread over the
lanes and the other indices looped, and per point the statements in order, each value in a
``const auto`` local, loads and stores straight to global memory.

gfx950 design. A pointwise kernel moves bytes and does little arithmetic, so the work is laid out
for HBM streaming instead of per element:

* the flat point space of ``EPB`` consecutive elements (``EPB * P`` points, ``P`` the points of
  one element, first index fastest) is spread over a 256-lane workgroup, ``ITEMS`` points per lane
  ``256`` apart, so consecutive lanes touch consecutive entries of a full-box tensor and one
  wavefront load covers 512 contiguous bytes (FP64); ``EPB`` is chosen so the ``256 * ITEMS``
  lanes are nearly all busy (``P = 504``: 2 elements, 1008 of 1024 lane-points);
* the point program runs in three phases per lane: every entry read by any statement is loaded
  for all ``ITEMS`` points first (all loads in flight at once), the statements then run on
  registers (a tensor entry written and read again is forwarded, not re-read), and the written
  entries are stored last;
* per-element entries whose first index runs along the unit-stride dimension are touched once
  in whole cache lines, so they are loaded and stored non-temporally (measured on the update
  ``Q += dt * I``: 6.0 TB/s against 5.4 TB/s with ordinary loads); shared operands and other
  access patterns keep ordinary, cached accesses;
* ``P > 256 * ITEMS`` loops over ``256 * ITEMS`` point chunks of one element.

The three-phase form assumes what the reference's per-lane form already assumes: points are
independent. For every tensor that is written, no two accesses (two points, or two index mappings
at one point) may reach the same entry; that is checked here, entry by entry.

Types follow C++ on the device: tensor entries have the tensor's datatype (the kernel's floating
type when unset), arithmetic promotes as in C++, math functions are the kernel type's
(``sqrt``/``sqrtf`` ...), comparisons give ``bool``. A literal written as a Python ``float`` takes
the kernel's floating type; the reference writes it as a single-precision literal (``0.1f``) even
in a double kernel (``optree.py:604-606``, ``basic_types.py:95-97``).
"""
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from ..common.basic_types import Addressing, FloatingPointType, GeneralLexicon
from ..common.exceptions import GenerationError
from ..common.operation import ARITHMETIC, BINARY_MATH, COMPARISON, UNARY_MATH, Operation
from ..generators import optree as ot
from .ir import SRC_DATA, SRC_GLOBAL, SRC_SCALAR, build_args, colmajor_strides
from .launcher import Emitted, Writer, kernel_params

THREADS = 256
ITEMS = 4

F = FloatingPointType
_INT_TYPES = {F.BOOL: (0, True), F.INT: (4, True), F.UINT: (4, False), F.LONG: (8, True),
              F.ULONG: (8, False)}


def promote(a, b):
  """C++ usual arithmetic conversions over the supported types."""
  if F.DOUBLE in (a, b):
    return F.DOUBLE
  if F.FLOAT in (a, b):
    return F.FLOAT
  a = F.INT if a == F.BOOL else a
  b = F.INT if b == F.BOOL else b
  if a == b:
    return a
  (sa, signed_a), (sb, signed_b) = _INT_TYPES[a], _INT_TYPES[b]
  if signed_a == signed_b:
    return a if sa >= sb else b
  u, s = (a, b) if not signed_a else (b, a)
  if _INT_TYPES[u][0] >= _INT_TYPES[s][0]:
    return u
  return s       # a wider signed type holds every value of the narrower unsigned one


def is_integer(t):
  return t in _INT_TYPES


@dataclass
class Cell:
  """One (tensor, index mapping) accessed by the statements."""
  index: int
  sub: object
  indices: Tuple[int, ...]
  arg: object
  dtype: FloatingPointType
  read: bool = False
  written: bool = False
  load: bool = False        # needs a load before the statements run
  defined: bool = False     # written unconditionally before any read

  @property
  def tensor(self):
    return self.sub.tensor


@dataclass
class PointwiseProgram:
  args: list
  statements: list
  extents: List[int]              # range of n0, n1, ... (bbox-relative)
  cells: List[Cell]
  fp_type: FloatingPointType
  cell_of: Dict[int, Cell] = field(default_factory=dict)   # id(TensorVar) -> cell
  temps: list = field(default_factory=list)

  @property
  def points(self):
    p = 1
    for x in self.extents:
      p *= x
    return p

  def flops_per_element(self):
    return 0


def _cell_offsets(cell, extents):
  """Entry offset (inside one element) of ``cell`` at every point, first index fastest."""
  t = cell.tensor
  strides = colmajor_strides(t.shape)
  rank = len(extents)
  grid = np.indices(list(reversed(extents)), dtype=np.int64) if rank else None
  off = np.zeros(int(np.prod(extents)) if rank else 1, dtype=np.int64)
  for dim, idx in enumerate(cell.indices):
    n = grid[rank - 1 - (-idx - 1)].reshape(-1)
    off += (n + cell.sub.bbox.lower()[dim] - t.bbox.lower()[dim]) * strides[dim]
  return off


def _type_of_tensor(sub, fp):
  dt = sub.tensor.datatype
  return fp if dt is None else dt


def lower(descr_list, global_tensors, fp_type) -> PointwiseProgram:
  """Statements, index extents and accessed entries of a list of ``ElementwiseDescr``."""
  args = build_args(global_tensors, fp_type)
  by_id = {id(a.tensor): a for a in args}
  statements, ranges = [], None
  for d in descr_list:
    r = {}
    for s in d.oplist:
      r = s.getRanges(r)
    if ranges is not None and r != ranges:
      raise GenerationError('ElementwiseDescrs of one kernel must range over the same index '
                            f'points: {ranges} vs {r}')
    ranges = r
    statements += d.oplist
  ranges = ranges or {}
  rank = len(ranges)
  if sorted(ranges) != [-i - 1 for i in reversed(range(rank))]:
    raise GenerationError(f'index ids must be -1..-{rank}, got {sorted(ranges)}')
  extents = [int(ranges[-i - 1][1]) for i in range(rank)]

  prog = PointwiseProgram(args=args, statements=statements, extents=extents, cells=[],
                          fp_type=fp_type)
  keyed = {}
  mapping_of_tensor = {}

  def visit_var(var, write):
    if isinstance(var, ot.TensorVar):
      t = var.tensor.tensor
      if t.is_tmp:
        raise GenerationError(f'{t.name}: chain temporaries cannot be used pointwise')
      if t.is_sparse():
        raise GenerationError(f'{t.name}: sparse tensors are not supported pointwise')
      arg = by_id.get(id(t))
      if arg is None:
        raise GenerationError(f'{t.name} is not in the kernel scope')
      if arg.kind == SRC_SCALAR:
        raise GenerationError(f'{t.name} is a by-value scalar: use scalar(), not tensor()')
      if arg.kind == SRC_DATA and write:
        raise GenerationError(f'{t.name} holds constant values and cannot be assigned to')
      key = (id(t), tuple(var.indices), tuple(var.tensor.bbox.lower()))
      cell = keyed.get(key)
      if cell is None:
        cell = Cell(index=len(prog.cells), sub=var.tensor, indices=tuple(var.indices), arg=arg,
                    dtype=_type_of_tensor(var.tensor, fp_type))
        keyed[key] = cell
        prog.cells.append(cell)
        mapping_of_tensor.setdefault(id(t), set()).add(key)
      prog.cell_of[id(var)] = cell
      if write:
        cell.written = True
      else:
        cell.read = True
    elif isinstance(var, ot.ScalarVar):
      if write:
        raise GenerationError('a scalar argument cannot be assigned to')
      t = var.tensor.tensor
      arg = by_id.get(id(t))
      if arg is None and not t.has_values():
        raise GenerationError(f'{t.name} is not in the kernel scope')
    elif isinstance(var, ot.TempVar):
      if all(v is not var for v in prog.temps):
        prog.temps.append(var)

  def visit_expr(node):
    for n in ot.walk_nodes(node):
      if isinstance(n, (ot.TensorVar, ot.ScalarVar, ot.TempVar)):
        visit_var(n, write=False)

  # program order: decides which cells must be loaded (read before any unconditional write)
  def scan(stmts, definite):
    for s in stmts:
      if isinstance(s, ot.Assignment):
        visit_expr(s.optree)
        visit_var(s.dest, write=True)
        if isinstance(s.dest, ot.TensorVar):
          cell = prog.cell_of[id(s.dest)]
          if definite and not cell.read:
            cell.defined = True
      elif isinstance(s, (ot.IfNode, ot.WhileNode)):
        visit_expr(s.condition)
        scan(s.subassignments, False)
      else:
        raise GenerationError(f'unknown statement {type(s).__name__}')

  scan(statements, True)
  for c in prog.cells:
    c.load = c.read and not c.defined
  for keys in mapping_of_tensor.values():
    cells = [keyed[k] for k in sorted(keys)]
    if any(c.written for c in cells):
      offsets = np.concatenate([_cell_offsets(c, extents) for c in cells])
      if len(np.unique(offsets)) != len(offsets):
        raise GenerationError(f'{cells[0].tensor.name} is written, and two accesses of it (at '
                              'one point or at two points) reach the same entry: the points '
                              'would depend on each other')
  return prog


class _Code:
  """Expression and statement code for one point (one item of a lane)."""

  def __init__(self, prog: PointwiseProgram, w: Writer, suffix: str):
    self.prog = prog
    self.w = w
    self.sfx = suffix
    self.count = 0

  def ctype(self, t):
    return str(t)

  def fresh(self):
    self.count += 1
    return f'v{self.count}{self.sfx}'

  def literal(self, node: ot.Immediate):
    t = self.imm_type(node)
    return t.literal(node.value)

  def imm_type(self, node):
    return self.prog.fp_type if node.fptype == F.FLOAT else node.fptype

  def type_of(self, node):
    fp = self.prog.fp_type
    if isinstance(node, ot.TensorVar):
      return self.prog.cell_of[id(node)].dtype
    if isinstance(node, ot.ScalarVar):
      return _type_of_tensor(node.tensor, fp)
    if isinstance(node, ot.TempVar):
      return fp if node.datatype is None else node.datatype
    if isinstance(node, ot.Immediate):
      return self.imm_type(node)
    if isinstance(node, ot.CastOpNode):
      return node.targetType
    if isinstance(node, ot.ConditionalOpNode):
      return promote(self.type_of(node.operands[1]), self.type_of(node.operands[2]))
    if isinstance(node, ot.LexicOpNode):
      op = node.optype
      ts = [self.type_of(o) for o in node.operands]
      if op == Operation.COPY:
        return ts[0]
      if op in ARITHMETIC:
        return promote(ts[0], ts[1])
      if op == Operation.NEG:
        return promote(ts[0], F.INT) if is_integer(ts[0]) else ts[0]
      if op == Operation.RCP:
        return promote(F.INT, ts[0])
      if op in UNARY_MATH or op in BINARY_MATH:
        return fp
      if op in COMPARISON:
        return F.BOOL
      if op in (Operation.AND, Operation.OR, Operation.XOR):
        if ts[0] == F.BOOL and ts[1] == F.BOOL:
          return F.BOOL
        return promote(ts[0], ts[1])
      if op == Operation.NOT:
        return F.BOOL if ts[0] == F.BOOL else promote(ts[0], F.INT)
    raise GenerationError(f'cannot type {node}')

  def expr(self, node) -> str:
    """C expression for ``node``; intermediate results go to ``const`` locals."""
    fp = self.prog.fp_type
    if isinstance(node, ot.TensorVar):
      return f'c{self.prog.cell_of[id(node)].index}{self.sfx}'
    if isinstance(node, ot.ScalarVar):
      t = node.tensor.tensor
      if t.has_values():
        vals = t.get_values()
        v = vals[()] if isinstance(vals, dict) else vals[0]
        return _type_of_tensor(node.tensor, fp).literal(v)
      return t.name
    if isinstance(node, ot.TempVar):
      return f'{node.name}{self.sfx}'
    if isinstance(node, ot.Immediate):
      return self.literal(node)
    vs = [self.expr(o) for o in node.operands]
    if isinstance(node, ot.CastOpNode):
      code = f'static_cast<{node.targetType}>({vs[0]})'
    elif isinstance(node, ot.ConditionalOpNode):
      code = f'(({vs[0]}) ? ({vs[1]}) : ({vs[2]}))'
    else:
      code = self.operation(node, vs)
    name = self.fresh()
    self.w(f'const {self.type_of(node)} {name} = {code};')
    return name

  def operation(self, node, v):
    op = node.optype
    fp = self.prog.fp_type
    f = 'f' if fp == F.FLOAT else ''
    ts = [self.type_of(o) for o in node.operands]
    simple = {Operation.ADD: '+', Operation.SUB: '-', Operation.MUL: '*', Operation.DIV: '/',
              Operation.LT: '<', Operation.LE: '<=', Operation.GT: '>', Operation.GE: '>=',
              Operation.EQ: '==', Operation.NEQ: '!='}
    if op == Operation.COPY:
      return v[0]
    if op in simple:
      return f'({v[0]} {simple[op]} {v[1]})'
    if op == Operation.MOD:
      if is_integer(ts[0]) and is_integer(ts[1]):
        return f'({v[0]} % {v[1]})'
      return f'fmod{f}({v[0]}, {v[1]})'
    if op == Operation.NEG:
      return f'(-{v[0]})'
    if op == Operation.RCP:
      return f'(1 / {v[0]})'
    if op in (Operation.AND, Operation.OR, Operation.XOR):
      if ts[0] == F.BOOL and ts[1] == F.BOOL:
        sym = {Operation.AND: '&&', Operation.OR: '||', Operation.XOR: '!='}[op]
      elif is_integer(ts[0]) and is_integer(ts[1]):
        sym = {Operation.AND: '&', Operation.OR: '|', Operation.XOR: '^'}[op]
      else:
        raise GenerationError(f'{op} needs integer or bool operands, got {ts}')
      return f'({v[0]} {sym} {v[1]})'
    if op == Operation.NOT:
      if ts[0] == F.BOOL:
        return f'(!{v[0]})'
      if is_integer(ts[0]):
        return f'(~{v[0]})'
      raise GenerationError('bitnot needs an integer or bool operand')
    if op in BINARY_MATH:
      fn = {Operation.MIN: 'fmin', Operation.MAX: 'fmax', Operation.POW: 'pow'}[op]
      return f'{fn}{f}({v[0]}, {v[1]})'
    if op == Operation.SIGN:
      return f'static_cast<{fp}>(({v[0]} > 0) - ({v[0]} < 0))'
    if op in UNARY_MATH:
      fn = {Operation.SQRT: 'sqrt', Operation.CBRT: 'cbrt', Operation.RSQRT: 'rsqrt',
            Operation.RCBRT: 'rcbrt', Operation.ABS: 'fabs', Operation.CEIL: 'ceil',
            Operation.FLOOR: 'floor', Operation.ROUND: 'round', Operation.TRUNC: 'trunc',
            Operation.LOGP1: 'log1p', Operation.EXPM1: 'expm1', Operation.GAMMA: 'tgamma',
            Operation.ERF: 'erf', Operation.LOG: 'log', Operation.EXP: 'exp',
            Operation.SIN: 'sin', Operation.COS: 'cos', Operation.TAN: 'tan',
            Operation.ASIN: 'asin', Operation.ACOS: 'acos', Operation.ATAN: 'atan',
            Operation.SINH: 'sinh', Operation.COSH: 'cosh', Operation.TANH: 'tanh',
            Operation.ASINH: 'asinh', Operation.ACOSH: 'acosh', Operation.ATANH: 'atanh'}[op]
      return f'{fn}{f}({v[0]})'
    raise GenerationError(f'operation {op} is not supported pointwise')

  def statements(self, stmts):
    for s in stmts:
      if isinstance(s, ot.Assignment):
        value = self.expr(s.optree)
        if isinstance(s.dest, ot.TensorVar):
          cell = self.prog.cell_of[id(s.dest)]
          self.w(f'c{cell.index}{self.sfx} = static_cast<{cell.dtype}>({value});')
          self.w(f'w{cell.index}{self.sfx} = true;')
        else:
          dt = self.type_of(s.dest)
          self.w(f'{s.dest.name}{self.sfx} = static_cast<{dt}>({value});')
      elif isinstance(s, ot.IfNode):
        cond = self.expr(s.condition)
        with self.w.block(f'if ({cond})'):
          self.statements(s.subassignments)
      elif isinstance(s, ot.WhileNode):
        with self.w.block('while (true)'):
          cond = self.expr(s.condition)
          self.w(f'if (!({cond})) break;')
          self.statements(s.subassignments)


def _address(cell: Cell, e: str, n: List[str]) -> str:
  """C expression of the entry of ``cell`` at point ``n`` of element ``e``."""
  t = cell.tensor
  arg = cell.arg
  strides = colmajor_strides(t.shape)
  origin = t.bbox.lower()
  const = sum((lo - o) * s for lo, o, s in zip(cell.sub.bbox.lower(), origin, strides))
  terms = []
  for dim, idx in enumerate(cell.indices):
    d = -idx - 1
    terms.append(n[d] if strides[dim] == 1 else f'{n[d]} * {strides[dim]}u')
  inner = ' + '.join([str(const)] + terms) if terms else str(const)
  name = arg.name
  off = f'{name}{GeneralLexicon.EXTRA_OFFSET}'
  if arg.kind == SRC_DATA:
    return f'{name}_values[{inner}]'
  if arg.addressing == Addressing.STRIDED:
    return f'{name}[{e} * {t.get_real_volume()}ull + {off} + ({inner})]'
  if arg.addressing == Addressing.PTR_BASED:
    return f'{name}[{e}][{off} + ({inner})]'
  return f'{name}[{inner}]'


def _streamed(cell: Cell) -> bool:
  """A per-element entry that consecutive lanes read at consecutive addresses (the first index
  runs along the tensor's first, unit-stride dimension): each wavefront access covers whole cache
  lines, read once, so it bypasses the caches. Other patterns rely on the caches to reuse the
  rest of a line and keep ordinary accesses."""
  return cell.arg.per_element and len(cell.indices) > 0 and cell.indices[0] == -1


def _tables(prog, w):
  """Constant-valued tensors read at run-time indices live in device tables."""
  done = set()
  for c in prog.cells:
    if c.arg.kind != SRC_DATA or c.arg.name in done:
      continue
    done.add(c.arg.name)
    t = c.tensor
    vol = t.get_real_volume()
    vals = [0.0] * vol
    strides = colmajor_strides(t.shape)
    from .ir import iterate_box
    for idx in iterate_box([(0, s) for s in t.shape]):
      v = t.value(idx)
      if v is not None:
        pos = sum((i - o) * s for i, o, s in zip(idx, t.bbox.lower(), strides))
        if 0 <= pos < vol:
          vals[pos] = v
    body = ', '.join(c.dtype.literal(v) for v in vals)
    w(f'static __device__ const {c.dtype} {c.arg.name}_values[{vol}] = {{{body}}};')


def plan(points, entry_bytes=8):
  """(elements per block, items per lane) for ``points`` points per element; entries narrower
  than 8 bytes get proportionally more points per lane, so a lane keeps as many bytes in
  flight."""
  items_max = ITEMS * max(1, 8 // max(1, entry_bytes))
  chunk = THREADS * items_max
  if points == 0:
    return 64, 1
  if points >= chunk:
    return 1, items_max
  epb = chunk // points
  items = -(-(epb * points) // THREADS)
  return epb, items


def emit(prog: PointwiseProgram, context, base_name: str) -> Emitted:
  kname = base_name
  P = prog.points
  widths = [c.dtype.size() for c in prog.cells if c.arg.per_element]
  epb, items = plan(P, max(widths) if widths else 8)
  PB = epb * P
  w = Writer()
  _tables(prog, w)
  params = ', '.join(kernel_params(prog, context))
  w(f'extern "C" __global__ void __launch_bounds__({THREADS}) {kname}({params})')
  with w.block(''):
    if P == 0:
      w('return;')
    else:
      w(f'const size_t e0 = static_cast<size_t>(blockIdx.x) * {epb}u;')
      loop = PB > THREADS * items
      head = (f'for (unsigned t0 = 0; t0 < {PB}u; t0 += {THREADS * items}u)' if loop
              else 'const unsigned t0 = 0;')
      if loop:
        with w.block(head):
          _chunk(prog, w, items, P, PB, epb)
      else:
        w(head)
        _chunk(prog, w, items, P, PB, epb)
  loads = [c for c in prog.cells if c.load]
  stores = [c for c in prog.cells if c.written]
  per_el = lambda c: c.arg.per_element           # noqa: E731
  load_bytes = P * sum(c.dtype.size() for c in loads if per_el(c))
  store_bytes = P * sum(c.dtype.size() for c in stores if per_el(c))
  bpe = load_bytes + store_bytes
  shared = sum(c.dtype.size() * P for c in loads if not per_el(c) and c.arg.kind == SRC_GLOBAL)
  return Emitted(kernel_name=kname, kernel_src=w.src(), block=(THREADS, 1, 1),
                 elements_per_block=epb, lds_bytes=0, emitter='pointwise',
                 schedule={'kind': 'pointwise', 'points': P, 'elements_per_block': epb,
                           'items': items, 'threads': THREADS,
                           'loads': len(loads), 'stores': len(stores),
                           'load_bytes_per_element': load_bytes,
                           'store_bytes_per_element': store_bytes},
                 bytes_per_element=bpe, shared_bytes=shared)


def _chunk(prog, w, items, P, PB, epb):
  rank = len(prog.extents)
  for k in range(items):
    s = f'_{k}'
    w(f'const unsigned t{s} = t0 + {k * THREADS}u + threadIdx.x;')
    w(f'const unsigned el{s} = t{s} / {P}u;')
    w(f'const size_t e{s} = e0 + el{s};')
    w(f'bool ok{s} = t{s} < {PB}u && e{s} < {GeneralLexicon.NUM_ELEMENTS};')
    w(f'if (ok{s} && {GeneralLexicon.FLAGS_NAME} != nullptr) '
      f'ok{s} = {GeneralLexicon.FLAGS_NAME}[e{s}] != 0u;')
    if rank:
      w(f'unsigned r{s} = t{s} - el{s} * {P}u;')
    for d, ext in enumerate(prog.extents):
      if d == rank - 1:
        w(f'const unsigned n{d}{s} = r{s};')
      else:
        w(f'const unsigned n{d}{s} = r{s} % {ext}u; r{s} /= {ext}u;')
  # phase 1: loads of every item
  for k in range(items):
    s = f'_{k}'
    n = [f'n{d}{s}' for d in range(rank)]
    for c in prog.cells:
      w(f'{c.dtype} c{c.index}{s} = {c.dtype.literal(0)};')
      if c.written:
        w(f'bool w{c.index}{s} = false;')
      if c.load:
        addr = _address(c, "e" + s, n)
        if _streamed(c):
          # entries of one element are read once: do not keep them in the caches
          w(f'if (ok{s}) c{c.index}{s} = __builtin_nontemporal_load(&{addr});')
        else:
          w(f'if (ok{s}) c{c.index}{s} = {addr};')
  # phase 2: the statements on registers
  for k in range(items):
    s = f'_{k}'
    with w.block(f'if (ok{s})'):
      for t in prog.temps:
        dt = prog.fp_type if t.datatype is None else t.datatype
        w(f'{dt} {t.name}{s} = {dt.literal(0)};')
      _Code(prog, w, s).statements(prog.statements)
  # phase 3: stores
  for k in range(items):
    s = f'_{k}'
    n = [f'n{d}{s}' for d in range(rank)]
    for c in prog.cells:
      if c.written:
        addr = _address(c, "e" + s, n)
        if _streamed(c):
          w(f'if (ok{s} && w{c.index}{s}) __builtin_nontemporal_store(c{c.index}{s}, &{addr});')
        else:
          w(f'if (ok{s} && w{c.index}{s}) {addr} = c{c.index}{s};')


def check_epilogue(prog: PointwiseProgram, dest, rows, cols):
  """Statements fuse into a chain's kernel when they range over exactly the rows ``rows`` and
  columns ``cols`` the chain produced, write nothing but the chain's final destination at the
  entries it wrote (index ids -1, -2), and read other tensors only through arguments (a point
  is a destination row of the lane's column); scalars and temporaries are free."""
  if prog.extents != [rows[1] - rows[0], cols[1] - cols[0]]:
    raise GenerationError(f'fused statements range over {prog.extents}, the chain writes '
                          f'{rows[1] - rows[0]} x {cols[1] - cols[0]}')
  for c in prog.cells:
    if c.tensor is dest:
      if c.indices != (-1, -2) or tuple(c.sub.bbox.lower()) != (rows[0], cols[0]):
        raise GenerationError('fused statements access the chain\'s destination only at the '
                              'entries it wrote')
    elif c.written:
      raise GenerationError('fused statements write nothing but the chain\'s destination')
    elif c.arg.kind != SRC_GLOBAL:
      raise GenerationError(f'{c.tensor.name}: constant tables are not read by fused statements')


def epilogue_address(cell: Cell, n) -> str:
  """Offset of ``cell``'s entry at point ``n`` (C expressions) from its element's base."""
  t = cell.tensor
  strides = colmajor_strides(t.shape)
  terms = [str(sum((lo - o) * st for lo, o, st in
                   zip(cell.sub.bbox.lower(), t.bbox.lower(), strides)))]
  for dim, idx in enumerate(cell.indices):
    terms.append(f'({n[-idx - 1]}) * {strides[dim]}')
  return ' + '.join(terms)
