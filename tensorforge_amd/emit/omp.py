"""OpenMP (host C++) dialect of a chain: the CPU form of the generated launcher.

The reference emits the same kernel body for CPUs through its ``omptarget`` backend
(``common/vm/lexic/target_lexic.py:39-205``): one OpenMP team per block of elements, one thread per
lane, LDS as team-shared arrays, barriers between phases. What fixes its results is the order of
the arithmetic for each output entry (``backend/instructions/compute/multilinear.py:105-157,
226-247``): the entry starts from zero; the contraction runs over the linearised contraction
indices (the first index fastest, ``backend/symbol.py:223-249``); each term is the product of the
operands in the description's order (``prod = prod * data``), added to the entry
(``value + prod``); an accumulating product then adds the destination's previous value
(``value + oldvalue``). This emitter keeps exactly that order per entry, so for finite inputs its
results are bit for bit the reference's (a sparse operand's missing entry contributes a zero
product there and nothing here). The thread structure is a plain ``omp parallel for`` over
elements, since the lanes of a team compute disjoint entries.

The launcher has the reference's name and parameter list (host pointers) and runs to completion
before it returns; ``flags[e] == 0`` skips element e. ``streamPtr`` is not read, but a null one
throws ``std::invalid_argument`` as the reference's launcher does
(``common/vm/lexic/target_lexic.py:179-184``). Semantics beyond the reference follow
this package's (a ``GemmDescr``'s ``alpha`` is a scalar operand of the product and its ``beta``
scales the previous value; permutes are honoured). Build with ``-O2 -fopenmp -ffp-contract=off`` (no fused multiply-add), as
``runtime.build.compile_host_source`` does.
"""
from ..common.basic_types import Addressing, FloatingPointType, GeneralLexicon
from ..common.exceptions import GenerationError
from . import ir
from .launcher import Writer

HEADER = '#include <cstddef>\n#include <cstdint>\n#include <stdexcept>\n'


def _ordered_operands(p: ir.Product):
  """The product's operands in the description's order (factors and scalar operands
  interleaved as the description lists them)."""
  if p.descr is None:
    return list(p.factors) + list(p.scalars)
  scalar_ids = {id(s.tensor) for s in p.scalars}
  fi = si = 0
  out = []
  for op in p.descr.ops:
    if id(op.tensor) in scalar_ids and si < len(p.scalars) and p.scalars[si].tensor is op.tensor:
      out.append(p.scalars[si])
      si += 1
    else:
      out.append(p.factors[fi])
      fi += 1
  return out


class _Omp:
  def __init__(self, program: ir.ChainProgram, generator):
    self.prog = program
    self.gen = generator
    self.real = FloatingPointType.as_str(program.fp_type)
    self.w = Writer()
    self.tables = []          # file-scope lookup tables (sparse positions, baked values)
    self.tmp_names = {}

  def lit(self, v):
    return self.prog.fp_type.literal(v)

  def emit(self):
    if self.prog.epilogue is not None:
      raise GenerationError('the OpenMP dialect generates multilinear chains only')
    w = self.w
    body = Writer()
    self.w = body
    body._depth = 2
    for a in self.prog.args:
      if a.kind != ir.SRC_GLOBAL:
        continue
      ty = f'{"const " if a.const else ""}{a.dtype}'
      if a.addressing == Addressing.STRIDED:
        body(f'{ty}* g_{a.name} = {a.name} + e * {a.element_stride}ull + {a.name}_extraOffset;')
      elif a.addressing == Addressing.PTR_BASED:
        body(f'{ty}* g_{a.name} = {a.name}[e] + {a.name}_extraOffset;')
      else:
        body(f'{ty}* g_{a.name} = {a.name};')
    for p in self.prog.products:
      if p.dest_src == ir.SRC_TMP and id(p.dest) not in self.tmp_names:
        name = f't{len(self.tmp_names)}'
        self.tmp_names[id(p.dest)] = name
        size = 1
        for s in self.prog.tmp_shapes[id(p.dest)]:
          size *= max(1, s)
        body(f'{self.real} {name}[{size}];')
    for i, p in enumerate(self.prog.products):
      self.product(i, p)
    self.w = w
    proto = self.gen._generate_launcher_proto(with_defaults=False)
    w.lines(HEADER)
    for t in self.tables:
      w(t)
    with w.block(f'extern "C" {proto}'):
      with w.block(f'if ({GeneralLexicon.STREAM_PTR_STR} == nullptr)'):
        w('throw std::invalid_argument("stream may not be null!");')
      w(f'const long long tf_n = static_cast<long long>({GeneralLexicon.NUM_ELEMENTS});')
      w('#pragma omp parallel for schedule(static)')
      with w.block('for (long long e = 0; e < tf_n; ++e)'):
        w(f'if ({GeneralLexicon.FLAGS_NAME} != nullptr && {GeneralLexicon.FLAGS_NAME}[e] == 0u) '
          'continue;')
        for line in body._lines:
          w._lines.append(line)
    return w.src()

  def access(self, f: ir.Factor, idx_exprs, name_hint):
    """(C expression of the factor's value at index ``idx_exprs``, or None for a product that
    is structurally zero there; a guard expression or None)."""
    if f.src == ir.SRC_SCALAR:
      return f.arg.name, None
    if f.src == ir.SRC_DATA:
      shape = [s for s in f.tensor.shape]
      size = 1
      for s in shape:
        size *= s
      vals = [0.0] * size
      strides = ir.colmajor_strides(shape)
      for idx, v in f.values.items():
        vals[sum(i * s for i, s in zip(idx, strides))] = v
      tab = f'tf_{self.gen.get_base_name()}_{name_hint}'
      self.tables.append(f'static const {self.real} {tab}[{size}] = '
                         f'{{{", ".join(self.lit(v) for v in vals)}}};')
      off = ' + '.join(f'({e}) * {s}' for e, s in zip(idx_exprs, strides)) or '0'
      return f'{tab}[{off}]', None
    base = self.tmp_names[id(f.tensor)] if f.src == ir.SRC_TMP else f'g_{f.arg.name}'
    if f.is_sparse():
      shape = list(f.tensor.shape)
      strides = ir.colmajor_strides(shape)
      size = 1
      for s in shape:
        size *= s
      pos = [-1] * size
      for idx, p in f.packed.items():
        pos[sum(i * s for i, s in zip(idx, strides))] = p
      tab = f'tf_{self.gen.get_base_name()}_{name_hint}'
      self.tables.append(f'static const int {tab}[{size}] = {{{", ".join(map(str, pos))}}};')
      off = ' + '.join(f'({e}) * {s}' for e, s in zip(idx_exprs, strides)) or '0'
      return f'{base}[{tab}[{off}]]', f'{tab}[{off}] >= 0'
    off = ' + '.join(f'(({e}) - {o}) * {s}' for e, o, s in zip(idx_exprs, f.origin, f.strides))
    return f'{base}[{off or "0"}]', None

  def product(self, i, p: ir.Product):
    w = self.w
    real = self.real
    w(f'// product {i}: {"accumulate" if p.add else "overwrite"}')
    if p.out_empty():
      return
    nv = [f'n{d}' for d in range(len(p.n_ranges))]
    kv = [f'k{d}' for d in range(len(p.k_ranges))]
    depth = 0
    # output entries (any order: each is computed on its own)
    for d in reversed(range(len(p.n_ranges))):
      lo, hi = p.n_ranges[d]
      w(f'for (int {nv[d]} = {lo}; {nv[d]} < {hi}; ++{nv[d]}) {{')
      w._depth += 1
      depth += 1
    w(f'{real} value = {self.lit(0.0)};')
    if not p.is_empty():
      # contraction: the first index fastest (the reference's linearised loop)
      for d in reversed(range(len(p.k_ranges))):
        lo, hi = p.k_ranges[d]
        w(f'for (int {kv[d]} = {lo}; {kv[d]} < {hi}; ++{kv[d]}) {{')
        w._depth += 1
        depth += 1
      terms, guards = [], []
      for j, f in enumerate(_ordered_operands(p)):
        idx = [nv[x] if kind == 'n' else kv[x] for kind, x in f.labels]
        expr, guard = self.access(f, idx, f'p{i}_o{j}')
        terms.append(expr)
        if guard:
          guards.append(guard)
      prod = terms[0]
      for t in terms[1:]:
        prod = f'{prod} * {t}'
      if guards:
        w(f'if ({" && ".join(guards)}) value = value + {prod};')
      else:
        w(f'value = value + {prod};')
      for _ in p.k_ranges:
        w._depth -= 1
        w('}')
        depth -= 1
    dest_off = ' + '.join(f'(({n}) - {o}) * {s}' for n, o, s in
                          zip(nv, p.dest_origin, p.dest_strides)) or '0'
    dest = self.tmp_names[id(p.dest)] if p.dest_src == ir.SRC_TMP else f'g_{p.dest_arg.name}'
    if p.add:
      old = f'{dest}[{dest_off}]'
      w(f'value = value + {old if p.beta == 1.0 else f"{self.lit(p.beta)} * {old}"};')
    w(f'{dest}[{dest_off}] = value;')
    while depth > 0:
      w._depth -= 1
      w('}')
      depth -= 1


def emit(program: ir.ChainProgram, generator) -> str:
  """C++/OpenMP source of the chain's launcher (a complete translation unit)."""
  return _Omp(program, generator).emit()
