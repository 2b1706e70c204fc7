"""The plain emitter: the reference's thread mapping restated for gfx950.

This is both the general fallback (any rank, any placement, any sparsity) and the "plain
generated kernel" baseline the tuned emitter is measured against. It keeps the reference's
structure (``generators/generator.py:128-214``, ``backend/instructions/builders/
multilinear_builder.py:51-260``, ``backend/instructions/compute/multilinear.py:98-157,226-247``):

* a block is ``(lanes, elements)``; lanes per element and elements per block follow the reference
  heuristics (``descriptions.py:35-49``, ``generator.py:38-50,201-214``);
* output dim 0 is spread over lanes in chunks of ``lanes`` rows; the remaining output dims are a
  register array per lane;
* an operand whose storage dim 0 is output dim 0 is read straight from global memory; any other
  operand is first staged into the element's LDS region; chain temporaries stay in registers when
  consumed along their leading dim and go through LDS otherwise;
* each product is stored to its destination when it finishes; ``add`` re-reads the destination;
* sparse operands are resolved at generation time: structural zeros are skipped, and a sparse
  operand read along the lane dimension uses range checks over runs of its packed values
  (``backend/symbol.py:340-401``).

Unlike the reference, every barrier is executed by all threads of the block (the element guard
wraps memory phases, never a barrier), so a partial last block or a ``flags`` mask cannot hang.
"""
from ..common.basic_types import Addressing, FloatingPointType
from ..common.exceptions import GenerationError
from . import ir
from .launcher import Emitted, Writer, kernel_params
from .traffic import element_traffic


UNROLLED_FENCE_EVERY = 1


def _real(fp):
  return FloatingPointType.as_str(fp)


def _lit(fp, v):
  return fp.literal(v)


def emit(program: ir.ChainProgram, context, base_name, num_threads, policy_type) -> Emitted:
  if program.epilogue is not None:
    raise GenerationError('the plain emitter does not fuse pointwise statements into a chain')
  fp = program.fp_type
  real = _real(fp)
  lanes = max(1, num_threads)
  kname = f'kernel_{base_name}'

  # ---- placement and LDS layout per product ----
  plans = []
  lds_per_elem = 0
  tmp_in_regs = {}
  for pi, p in enumerate(program.products):
    plan = {'staged': {}}
    offset = 0
    for fi, f in enumerate(p.factors):
      direct = f.labels and f.labels[0] == ('n', 0)
      if f.src == ir.SRC_TMP:
        if direct and p.n_ranges[0][0] == 0:
          continue
        shape = program.tmp_shapes[id(f.tensor)]
        ld0 = shape[0]
        size = ld0 * ir._prod(shape[1:])
        plan['staged'][fi] = ('tmp', offset, (ld0,) + tuple(shape[1:]))
        offset += size
      elif f.src == ir.SRC_GLOBAL:
        if direct:
          continue
        if f.is_sparse():
          size = max(f.packed.values()) + 1 if f.packed else 0
          plan['staged'][fi] = ('sparse', offset, size)
          offset += size
        else:
          window = [_label_range(p, lab) for lab in f.labels]
          sizes = [hi - lo for lo, hi in window]
          ld0 = sizes[0]
          if any(lab == ('n', 0) for lab in f.labels[1:]) and ld0 % 2 == 0:
            ld0 += 1        # odd leading extent: lanes walking dim>0 spread over banks
          plan['staged'][fi] = ('dense', offset, window, [ld0] + sizes[1:])
          offset += ld0 * ir._prod(sizes[1:])
      offset = (offset + 1) // 2 * 2
    plan['lds'] = offset
    lds_per_elem = max(lds_per_elem, offset)
    plans.append(plan)
  if lds_per_elem > 0:
    # each element's region padded as the reference sizes it: aligned to the vector unit's width
    # in words, plus its bank term (reference common/context.py:38-44, backend/opt/optimizer.py:35-37)
    lds_per_elem = context.align(lds_per_elem) + (lanes % 32) // 4 * 4

  policy = policy_type(context, 0, lds_per_elem, lanes)
  mults = policy.get_num_mults_per_block()
  lds_bytes = lds_per_elem * mults * fp.size()
  if lds_bytes > context.get_hw().max_local_mem_size_per_block:
    raise GenerationError(f'plain kernel needs {lds_bytes} B of LDS per block')

  w = Writer()
  params = ', '.join(kernel_params(program, context))
  w(f'extern "C" __global__ void __launch_bounds__({lanes * mults}) {kname}({params})')
  w('{')
  w._depth += 1
  w(f'// plain emitter: {lanes} lanes/element, {mults} elements/block, '
    f'{lds_per_elem} LDS words/element')
  if lds_per_elem > 0:
    w(f'extern __shared__ __align__(16) {real} totalShrMem[];')
    w(f'{real}* __restrict__ lds = &totalShrMem[{lds_per_elem} * threadIdx.y];')
  w('const int lane = threadIdx.x;')
  w('const size_t batchId = threadIdx.y + (size_t)blockDim.y * blockIdx.x;')
  w('const bool active = (batchId < numElements) && (flags == nullptr || flags[batchId] != 0u);')
  for a in program.args:
    if a.kind != ir.SRC_GLOBAL:
      continue
    ty = f'{"const " if a.const else ""}{a.dtype}'
    if a.addressing == Addressing.NONE:
      w(f'{ty}* __restrict__ g_{a.name} = {a.name};')
    elif a.addressing == Addressing.STRIDED:
      w(f'{ty}* __restrict__ g_{a.name} = {a.name} + batchId * {a.element_stride}ull + '
        f'{a.name}_extraOffset;')
    else:
      w(f'{ty}* __restrict__ g_{a.name} = active ? {a.name}[batchId] + {a.name}_extraOffset '
        f': nullptr;')

  for pi, p in enumerate(program.products):
    _emit_product(w, program, p, pi, plans[pi], lanes, fp, tmp_in_regs)

  w._depth -= 1
  w('}')
  traffic = element_traffic(program)
  return Emitted(kernel_name=kname, kernel_src=w.src(), block=(lanes, mults, 1),
                 elements_per_block=mults, lds_bytes=lds_bytes, emitter='plain',
                 schedule={'lanes': lanes, 'elements_per_block': mults,
                           'lds_words_per_element': lds_per_elem},
                 bytes_per_element=traffic.per_element, shared_bytes=traffic.shared)


def _label_range(p, lab):
  kind, i = lab
  return p.n_ranges[i] if kind == 'n' else p.k_ranges[i]


def _tmp_reg_name(tensor):
  return f'reg_{tensor.name}'


def _chunks(size, lanes):
  return (size + lanes - 1) // lanes


def _emit_product(w, program, p, pi, plan, lanes, fp, tmp_in_regs):
  real = _real(fp)
  w()
  w(f'// product {pi}: {p.descr}')
  if p.out_empty():
    if p.dest_src == ir.SRC_TMP and id(p.dest) not in tmp_in_regs:
      raise GenerationError(f'temporary {p.dest.name} has an empty range')
    return
  n0_lo, n0_hi = p.n_ranges[0]
  n0 = n0_hi - n0_lo
  chunks = _chunks(n0, lanes)
  rest_ranges = p.n_ranges[1:]
  rest = ir._prod([hi - lo for lo, hi in rest_ranges])
  lds = 'lds'

  # ---- staging phase ----
  if plan['staged']:
    with w.block('if (active)'):
      for fi, info in plan['staged'].items():
        f = p.factors[fi]
        if info[0] == 'sparse':
          _, off, size = info
          w(f'for (int i = lane; i < {size}; i += {lanes}) '
            f'{lds}[{off} + i] = g_{f.arg.name}[i];')
        elif info[0] == 'dense':
          _, off, window, ldsizes = info
          sizes = [hi - lo for lo, hi in window]
          count = ir._prod(sizes)
          if count == 0:       # an empty contraction range stages nothing
            continue
          with w.block(f'for (int i = lane; i < {count}; i += {lanes})'):
            w('int r = i;')
            gidx, lidx, lstride = [], [], 1
            for d, (lo, hi) in enumerate(window):
              s = hi - lo
              w(f'const int i{d} = r % {s}; r /= {s};')
              gidx.append(f'({lo - f.origin[d]} + i{d}) * {f.strides[d]}')
              lidx.append(f'i{d} * {lstride}')
              lstride *= ldsizes[d]
            w(f'{lds}[{off} + {" + ".join(lidx)}] = g_{f.arg.name}[{" + ".join(gidx)}];')
        else:  # tmp registers -> LDS
          _, off, shape = info
          treg = _tmp_reg_name(f.tensor)
          tchunks = _chunks(shape[0], lanes)
          trest = shape[1:]
          for c in range(tchunks):
            guard = shape[0] - lanes * c
            body = []
            for r_idx, idx in enumerate(ir.iterate_box([(0, s) for s in trest])):
              lin = sum(i * s for i, s in zip(idx, ir.colmajor_strides(trest)))
              body.append(f'{lds}[{off} + lane + {lanes * c} + {shape[0] * lin}] = '
                          f'{treg}[{c + tchunks * r_idx}];')
            _guarded(w, guard, lanes, body)
  w('__syncthreads();')

  # ---- compute phase ----
  acc = f'acc{pi}'
  if p.dest_src == ir.SRC_TMP and id(p.dest) not in tmp_in_regs:
    shape = program.tmp_shapes[id(p.dest)]
    tchunks = _chunks(shape[0], lanes)
    trest = ir._prod(shape[1:])
    w(f'{real} {_tmp_reg_name(p.dest)}[{max(1, tchunks * trest)}];')
    tmp_in_regs[id(p.dest)] = (tchunks, trest)
  w(f'{real} {acc}[{chunks * rest}];')
  with w.block('if (active)'):
    for c in range(chunks):
      guard = n0 - lanes * c
      body = []
      for r_idx in range(rest):
        body.append(f'{acc}[{c + chunks * r_idx}] = {_lit(fp, 0)};')
      body.extend(_term_lines(p, plan, c, lanes, fp, acc, chunks, program.tmp_shapes,
                              tmp_in_regs))
      # previous value
      if p.add:
        for r_idx, nrest in enumerate(ir.iterate_box(rest_ranges)):
          prev = _dest_value(p, c, nrest, lanes, chunks, r_idx)
          scaled = prev if p.beta == 1.0 else f'{_lit(fp, p.beta)} * {prev}'
          body.append(f'{acc}[{c + chunks * r_idx}] = {acc}[{c + chunks * r_idx}] + {scaled};')
      # store
      for r_idx, nrest in enumerate(ir.iterate_box(rest_ranges)):
        if p.dest_src == ir.SRC_GLOBAL:
          off = p.dest_offset((n0_lo,) + nrest)
          body.append(f'g_{p.dest_arg.name}[lane + {lanes * c + off}] = {acc}[{c + chunks * r_idx}];')
        else:
          body.append(f'{_tmp_reg_name(p.dest)}[{c + chunks * r_idx}] = '
                      f'{acc}[{c + chunks * r_idx}];')
      _guarded(w, guard, lanes, body)
  w('__syncthreads();')


def _guarded(w, remaining, lanes, body):
  if not body:
    return
  if remaining >= lanes:
    with w.block(''):
      for line in body:
        w(line)
  else:
    with w.block(f'if (lane < {remaining})'):
      for line in body:
        w(line)


def _dest_value(p, c, nrest, lanes, chunks, r_idx):
  if p.dest_src == ir.SRC_GLOBAL:
    off = p.dest_offset((p.n_ranges[0][0],) + tuple(nrest))
    return f'g_{p.dest_arg.name}[lane + {lanes * c + off}]'
  return f'{_tmp_reg_name(p.dest)}[{c + chunks * r_idx}]'


def _affine(terms, const=0):
  """``const + sum(coef * var)`` for (coef, var) terms whose var is an int or a C expression."""
  parts = []
  for coef, var in terms:
    if isinstance(var, int):
      const += coef * var
    elif coef != 0:
      parts.append(f'{var} * {coef}' if coef != 1 else f'{var}')
  parts.append(str(const))
  return ' + '.join(parts)


def _runtime_k_ok(p, plan):
  """Whether the contraction can be a runtime loop: every factor is addressable with runtime
  contracted indices (no compile-time sparsity, no register temporaries indexed by k)."""
  if not p.k_ranges:
    return False
  for fi, f in enumerate(p.factors):
    has_k = any(kind == 'k' for kind, _ in f.labels)
    if f.src == ir.SRC_DATA and has_k:
      return False
    if f.src == ir.SRC_GLOBAL and f.is_sparse():
      return False
    if f.src == ir.SRC_TMP and fi not in plan['staged'] and has_k:
      return False
  return True


def _term_lines(p, plan, c, lanes, fp, acc, chunks, tmp_shapes, tmp_in_regs):
  """Accumulation statements for chunk ``c``: k outer, remaining output dims inner."""
  lines = []
  n0_lo = p.n_ranges[0][0]
  rest_ranges = p.n_ranges[1:]
  row = f'(lane + {lanes * c})'
  cache = {}
  counter = [0]

  def factor_expr(fi, f, n, k):
    """C expression for factor ``f`` at (n, k) where n[0] is the lane's row; None = zero.
    Entries of ``k`` are ints (unrolled) or C variable names (runtime loop)."""
    idx = ir.factor_index(f, n, k)
    lead_dims = [d for d, lab in enumerate(f.labels) if lab == ('n', 0)]
    if f.src == ir.SRC_DATA:
      v = f.values.get(idx)
      return None if not v else _lit(fp, v)
    if f.src == ir.SRC_SCALAR:
      return f.arg.name
    staged = plan['staged'].get(fi)
    if f.src == ir.SRC_TMP and staged is None:
      shape_rest = list(tmp_shapes[id(f.tensor)][1:])
      tmp_chunks = tmp_in_regs[id(f.tensor)][0]
      lin = sum(i * s for i, s in zip(idx[1:], ir.colmajor_strides(shape_rest)))
      return f'{_tmp_reg_name(f.tensor)}[{c + tmp_chunks * lin}]'
    if f.src == ir.SRC_TMP:
      _, off, shape = staged
      strides = ir.colmajor_strides(shape)
      terms = []
      for d, i in enumerate(idx):
        terms.append((strides[d], f'({row} + {n0_lo})' if d in lead_dims else i))
      return f'lds[{_affine(terms, off)}]'
    # global operand
    if f.is_sparse():
      if not lead_dims:
        pos = f.packed.get(idx)
        if pos is None:
          return None
        if staged is not None:
          return f'lds[{staged[1] + pos}]'
        return f'g_{f.arg.name}[{pos}]'
      key = (fi, tuple(x for d, x in enumerate(idx) if d not in lead_dims))
      if key not in cache:
        runs = _runs(f, idx, lead_dims, p.n_ranges[0])
        if not runs:
          cache[key] = None
        else:
          var = f'sv{counter[0]}'
          counter[0] += 1
          lines.append(f'{_real(fp)} {var} = {_lit(fp, 0)};')
          base = f'lds[{staged[1]} + ' if staged is not None else f'g_{f.arg.name}['
          for lo_row, hi_row, pos in runs:
            lines.append(f'if ({row} >= {lo_row} && {row} < {hi_row}) {var} = '
                         f'{base}{pos - lo_row} + {row}];')
          cache[key] = var
      return cache[key]
    if staged is not None:
      _, off, window, ldsizes = staged
      terms, stride, const = [], 1, off
      for d, i in enumerate(idx):
        if d in lead_dims:
          terms.append((stride, row))
          const += (n0_lo - window[d][0]) * stride
        else:
          terms.append((stride, i))
          const -= window[d][0] * stride
        stride *= ldsizes[d]
      return f'lds[{_affine(terms, const)}]'
    terms, const = [], 0
    for d, i in enumerate(idx):
      if d in lead_dims:
        terms.append((f.strides[d], row))
        const += (n0_lo - f.origin[d]) * f.strides[d]
      else:
        terms.append((f.strides[d], i))
        const -= f.origin[d] * f.strides[d]
    return f'g_{f.arg.name}[{_affine(terms, const)}]'

  def term_statements(k, out):
    for r_idx, nrest in enumerate(ir.iterate_box(rest_ranges)):
      n = (n0_lo,) + tuple(nrest)
      exprs = []
      zero = False
      for fi, f in enumerate(p.factors):
        e = factor_expr(fi, f, n, k)
        if e is None:
          zero = True
          break
        exprs.append(e)
      if zero:
        continue
      for s in p.scalars:
        if s.src == ir.SRC_DATA:
          v = next(iter(s.values.values()), None) if s.values else None
          if not v:
            zero = True
            break
          exprs.append(_lit(fp, v))
        else:
          exprs.append(s.arg.name)
      if zero:
        continue
      prod = ' * '.join(exprs) if exprs else _lit(fp, 1)
      a = f'{acc}[{c + chunks * r_idx}]'
      out.append(f'{a} = {a} + {prod};')

  if any(hi <= lo for lo, hi in p.k_ranges):
    pass                 # empty contraction: the accumulator stays zero
  elif _runtime_k_ok(p, plan):
    total = ir._prod([hi - lo for lo, hi in p.k_ranges])
    kvars = [f'k{i}' for i in range(len(p.k_ranges))]
    lines.append('#pragma unroll 2')
    lines.append(f'for (int kv = 0; kv < {total}; ++kv) {{')
    div = 1
    for i, (lo, hi) in enumerate(p.k_ranges):
      lines.append(f'  const int k{i} = {lo} + (kv / {div}) % {hi - lo};')
      div *= hi - lo
    body = []
    term_statements(kvars, body)
    lines.extend('  ' + b for b in body)
    lines.append('}')
  else:
    # Fully unrolled (compile-time sparsity). A scheduling fence every few contraction steps
    # keeps the compiler from hoisting every load of the unrolled chain ahead of its use,
    # which otherwise exhausts the register file and spills.
    for step, k in enumerate(ir.iterate_box(p.k_ranges)):
      if step and step % UNROLLED_FENCE_EVERY == 0:
        lines.append('__builtin_amdgcn_sched_barrier(0);')
      term_statements(k, lines)
  return lines


def _runs(f, idx, lead_dims, n0_range):
  """Runs ``(row_lo, row_hi, pos_at_row_lo)`` of stored entries along the lane dimension, rows
  relative to the product's first output row."""
  d = lead_dims[0]
  lo, hi = n0_range
  runs = []
  start = None
  for r in range(lo, hi):
    full = list(idx)
    for ld in lead_dims:
      full[ld] = r
    pos = f.packed.get(tuple(full))
    if pos is not None and start is not None and pos - start[1] == r - start[0]:
      continue
    if start is not None:
      runs.append((start[0] - lo, r - lo, start[1]))
      start = None
    if pos is not None:
      start = (r, pos)
  if start is not None:
    runs.append((start[0] - lo, hi - lo, start[1]))
  return runs
