"""Streaming schedule of the tuned emitter: persistent wavefronts, shared operands in LDS,
DPP-broadcast FMAs.

Applies to chains of the form (SeisSol's volume and local-flux chains)::

    D_e = [beta D_e]  +  sum_t  U_t · R_t,e            (one destination, one column per lane)
    R_t,e = L_e · W_t,e     (a temporary, formed on the fly)   or   R_t,e = B_t,e (an operand)

with every ``U_t`` shared by all elements (``Addressing.NONE``, dense or sparse) and ``L_e`` one
per-element operand shared by all temporaries.

Mapping. A workgroup of ``W`` wavefronts stays resident and walks element groups (``64 // q``
elements per wavefront). At start it copies the stored entries of every ``U_t`` into LDS in the
order the FMAs consume them. Per element group, the contraction runs in blocks of ``k_block``
rows: the block's rows of ``L_e`` are loaded one block ahead into registers and written to a
two-slot LDS ring; each lane forms its temporary's rows from the ring (a short per-lane list of
(column, weight) slots for a sparse ``W``), then every stored entry of ``U_t`` costs one
``v_fmac_f64_dpp`` whose shared operand comes from a 16-entry LDS chunk broadcast across each row
of 16 lanes (``row_newbcast``). No scalar loads and no lane branches are in the loop; the
destination column stays in registers and is written once per element.
"""
from dataclasses import dataclass, field
from typing import List, Optional

from ..common.basic_types import Addressing, FloatingPointType
from ..common.exceptions import GenerationError
from . import ir
from .launcher import Emitted, Writer, kernel_params, persistent_grid
from .traffic import element_traffic

WAVE = 64
CHUNK = 16


@dataclass
class Term:
  product: ir.Product
  U: ir.Factor
  u_dims: tuple                 # storage dims (row, k) of U
  kind: str                     # 'fused' or 'global'
  producer: Optional[ir.Product] = None
  W: Optional[ir.Factor] = None     # weights of a fused temporary (right operand of producer)
  w_dims: tuple = ()
  B: Optional[ir.Factor] = None     # per-element right operand of a 'global' term
  b_dims: tuple = ()
  entries: list = field(default_factory=list)   # (row rel, k, position in U)
  L: Optional[ir.Factor] = None     # streamed operand of a fused term, and its column window
  l_cols: tuple = ()
  l_shift: int = 0                  # row of L (temporary coordinates) that contraction index 0 reads


@dataclass
class Plan:
  dest: ir.GlobalArg
  dest_product: ir.Product
  rows: tuple
  cols: tuple
  q: int
  klo: int
  khi: int
  init: str                      # 'zero' or 'load'
  beta: float
  terms: List[Term]
  L: Optional[ir.Factor] = None  # streamed per-element operand of fused terms
  l_cols: tuple = ()             # column window of L
  # register stages after the streamed one: D = U_n · ( ... U_2 · (streamed stage)); the streamed
  # stage's accumulator is a temporary, the last stage writes the destination
  post: list = field(default_factory=list)
  final_init: str = 'zero'
  final_beta: float = 0.0
  # fused terms stream different operands (the time-derivative step that is not aligned): their
  # contraction blocks run one term after the other, each block's rows loaded from its term's
  # operand into the same ring
  stacked: bool = False


@dataclass
class Post:
  product: ir.Product
  U: ir.Factor
  u_dims: tuple
  rows: tuple                    # destination rows of the stage
  prev_lo: int                   # first row of the previous stage's accumulator
  entries: list = field(default_factory=list)   # (row rel, prev row rel, position in U)


def _classify(p):
  from .tuned import _classify as c
  return c(p)


def plan(program: ir.ChainProgram):
  """Recognise the streaming form; returns (Plan, None) or (None, reason)."""
  pl, why = _plan_single(program)
  if pl is not None:
    return pl, None
  chain, why2 = _plan_chain(program)
  if chain is not None:
    return chain, None
  return None, why


def _plan_chain(program):
  """``D (+)= U_n · (... (U_1 · (L · W)))``: shared left operands applied one after the other to
  a temporary formed from per-element operands (SeisSol's flux integrals). The innermost
  product streams like a single-stage chain; the outer ones run on the lane's registers."""
  prods = program.products
  consumers = [p for p in prods if p.dest_src == ir.SRC_GLOBAL]
  if len(consumers) != 1:
    return None, 'chain: not one destination product'
  final = consumers[0]
  by_dest = {id(q.dest): q for q in prods if q.dest_src == ir.SRC_TMP}
  stages = []
  cur = final
  while True:
    if cur.scalars:
      return None, 'chain: scaled product'
    sh, why = _classify(cur)
    if sh is None:
      return None, why
    if sh.left.src != ir.SRC_GLOBAL or sh.left.arg.addressing != Addressing.NONE:
      return None, 'chain: left operand is not shared'
    if sh.right.src != ir.SRC_TMP or sh.right_dims != (0, 1):
      return None, 'chain: right operand is not a temporary'
    q = by_dest.get(id(sh.right.tensor))
    if q is None or q.add or q.scalars:
      return None, 'chain: temporary not produced once'
    qsh, why = _classify(q)
    if qsh is None:
      return None, why
    stages.append((cur, sh))
    if qsh.left.src == ir.SRC_GLOBAL and qsh.left.arg.addressing == Addressing.NONE:
      cur = q
      continue
    inner, inner_sh, producer, psh = cur, sh, q, qsh
    break
  if len(stages) < 2:
    return None, 'chain: single stage'
  stages.reverse()              # innermost (streamed) first
  if len(stages) - 1 + 2 != len(prods):
    return None, 'chain: other products'
  if psh.left.src != ir.SRC_GLOBAL or not psh.left.arg.per_element or psh.left.is_sparse() or \
      psh.left_dims != (0, 1) or psh.right.src != ir.SRC_GLOBAL or not psh.right.arg.per_element:
    return None, 'chain: temporary not formed from per-element operands'
  rows, cols = tuple(inner.n_ranges[0]), tuple(inner.n_ranges[1])
  q = cols[1] - cols[0]
  if not 1 <= q <= WAVE // 2:
    return None, f'{q} columns'
  klo, khi = inner.k_ranges[0]
  if tuple(producer.n_ranges[0]) != (0, khi) or klo != 0 or producer.n_ranges[1][0] != 0 or \
      producer.n_ranges[1][1] < q:
    return None, 'chain: temporary rows or columns do not match'
  post = []
  prev_lo = rows[0]
  for p, sh in stages[1:]:
    if tuple(p.n_ranges[1]) != cols or p.n_ranges[0][1] - p.n_ranges[0][0] > 96:
      return None, 'chain: stage columns or rows'
    post.append(Post(p, sh.left, sh.left_dims, tuple(p.n_ranges[0]), prev_lo))
    prev_lo = p.n_ranges[0][0]
  if final.add and final.beta != 1.0:
    return None, 'chain: scaled accumulation'
  terms = [Term(inner, inner_sh.left, inner_sh.left_dims, 'fused', producer=producer,
                W=psh.right, w_dims=psh.right_dims)]
  return Plan(dest=final.dest_arg, dest_product=final, rows=rows, cols=cols, q=q, klo=klo,
              khi=khi, init='zero', beta=0.0, terms=terms, L=psh.left,
              l_cols=tuple(producer.k_ranges[0]), post=post,
              final_init='load' if final.add else 'zero',
              final_beta=final.beta if final.add else 0.0), None


def _plan_single(program):
  prods = program.products
  consumers = [p for p in prods if p.dest_src == ir.SRC_GLOBAL]
  producers = [p for p in prods if p.dest_src == ir.SRC_TMP]
  if not consumers:
    return None, 'no global destination'
  dest = consumers[0].dest_arg
  rows, cols = tuple(consumers[0].n_ranges[0]), tuple(consumers[0].n_ranges[1])
  q = cols[1] - cols[0]
  if not 1 <= q <= WAVE // 2:
    return None, f'{q} columns'
  if rows[1] - rows[0] > 96:
    return None, 'more than 96 rows'
  by_tmp = {}
  for p in producers:
    if p.add or p.scalars:
      return None, 'accumulating or scaled temporary'
    sh, why = _classify(p)
    if sh is None:
      return None, why
    if sh.left.src != ir.SRC_GLOBAL or not sh.left.arg.per_element or sh.left.is_sparse():
      return None, 'temporary whose left operand is not a dense per-element operand'
    if sh.left_dims != (0, 1):
      return None, 'transposed left operand of a temporary'
    if sh.right.src != ir.SRC_GLOBAL or not sh.right.arg.per_element:
      return None, 'temporary whose right operand is not per element'
    by_tmp[id(p.dest)] = (p, sh)
  terms = []
  L = None
  l_window = None
  klo = khi = None
  init, beta = 'zero', 0.0
  for n, p in enumerate(consumers):
    if p.dest_arg is not dest or tuple(p.n_ranges[0]) != rows or tuple(p.n_ranges[1]) != cols:
      return None, 'consumers write different destinations or windows'
    if p.scalars:
      return None, 'scaled product'
    if n == 0:
      if p.add:
        init, beta = 'load', p.beta
    elif not p.add or p.beta != 1.0:
      return None, 'a later product overwrites the destination'
    sh, why = _classify(p)
    if sh is None:
      return None, why
    U = sh.left
    if U.src != ir.SRC_GLOBAL or U.arg.addressing != Addressing.NONE:
      return None, 'left operand is not shared by all elements'
    # terms whose temporaries stream one operand may contract over different leading ranges
    # [0, k_t) of it (each term's entries stay in its own range); the plan's range is the widest
    kr = tuple(p.k_ranges[0])
    R = sh.right
    if klo is not None and kr != (klo, khi) and \
        (R.src != ir.SRC_TMP or kr[0] != 0 or klo != 0 or
         any(t.kind != 'fused' for t in terms)):
      return None, 'products with different contraction ranges'
    klo, khi = kr[0], max(kr[1], khi or 0)
    if R.src == ir.SRC_TMP:
      if id(R.tensor) not in by_tmp or sh.right_dims != (0, 1):
        return None, 'temporary not formed by a streamable product'
      prod, psh = by_tmp.pop(id(R.tensor))
      shift = prod.n_ranges[0][0]
      if prod.n_ranges[0][1] - shift < kr[1] or kr[0] != 0:
        return None, 'temporary rows do not match the contraction'
      if prod.n_ranges[1][0] != 0 or prod.n_ranges[1][1] < q:
        return None, 'temporary columns do not match'
      if L is None:
        L = psh.left
        l_window = tuple(prod.k_ranges[0])
      elif tuple(prod.k_ranges[0]) != l_window:
        return None, 'temporaries stream different column windows'
      terms.append(Term(p, U, sh.left_dims, 'fused', producer=prod, W=psh.right,
                        w_dims=psh.right_dims, L=psh.left, l_cols=tuple(prod.k_ranges[0]),
                        l_shift=shift))
    elif R.src == ir.SRC_GLOBAL and R.arg.per_element and not R.is_sparse():
      if any(tuple(t.product.k_ranges[0]) != kr for t in terms):
        return None, 'products with different contraction ranges'
      terms.append(Term(p, U, sh.left_dims, 'global', B=R, b_dims=sh.right_dims))
    else:
      return None, 'right operand kind'
  if by_tmp:
    return None, 'temporary used other than by a single streamed product'
  stacked = len({id(t.L.arg) for t in terms if t.kind == 'fused'}) > 1
  if not stacked and any(t.l_shift for t in terms if t.kind == 'fused'):
    return None, 'temporary rows start inside the streamed operand'
  if stacked:
    if any(t.kind != 'fused' for t in terms):
      return None, 'streamed operands of several temporaries beside a per-element operand'
    for t in terms:
      if t.L.arg.addressing != Addressing.STRIDED or t.L.strides != L.strides or \
          t.L.origin[1] != L.origin[1]:
        return None, 'temporaries stream operands of different layouts'
  # every product must be a consumer or its producer
  return Plan(dest=dest, dest_product=consumers[0], rows=rows, cols=cols, q=q, klo=klo, khi=khi,
              init=init, beta=beta, terms=terms, L=L, l_cols=l_window or (),
              stacked=stacked), None


class _Stream:
  def __init__(self, program, context, base_name, schedule, pl: Plan, allow_dma=True,
               suffix=''):
    self.allow_dma = allow_dma
    self.prog = program
    self.ctx = context
    self.fp = program.fp_type
    self.real = FloatingPointType.as_str(self.fp)
    self.sched = schedule
    self.pl = pl
    self.base_name = base_name
    self.kname = f'kernel_{base_name}{suffix}'
    self.w = Writer()
    self.q = pl.q
    self.G = WAVE // pl.q
    self.W = max(1, schedule.waves_per_block)
    self.KB = max(2, schedule.k_block)
    self.n = 0

  def fresh(self, p):
    self.n += 1
    return f'{p}{self.n}'

  def lit(self, v):
    return self.fp.literal(v)

  # ---------------------------------------------------------------- layout
  def order_entries(self):
    """Traversal order of every stored U entry: block, term, then k, row (columns of a block are
    visited one after the other so consecutive FMAs update different rows)."""
    pl = self.pl
    lo0 = pl.rows[0]
    seq = []
    for t_idx, t in enumerate(pl.terms):
      t.entries = []
      U = t.U
      nd, kd = t.u_dims
      if U.is_sparse():
        items = U.packed.items()
      else:
        items = [(idx, U.offset(idx)) for idx in
                 ir.iterate_box([(U.origin[d], U.origin[d] + U.tensor.shape[d]) for d in range(2)])]
      tk0, tk1 = t.product.k_ranges[0]
      for idx, pos in items:
        i, k = idx[nd], idx[kd]
        if pl.rows[0] <= i < pl.rows[1] and max(pl.klo, tk0) <= k < min(pl.khi, tk1):
          t.entries.append((i - lo0, k, pos))
    if pl.stacked:
      # one term after the other, each over its own contraction range
      blocks = []
      for t_idx, t in enumerate(pl.terms):
        tk0, tk1 = t.product.k_ranges[0]
        blocks += [(b0, min(b0 + self.KB, tk1), (t_idx,)) for b0 in range(tk0, tk1, self.KB)]
    else:
      blocks = [(b0, min(b0 + self.KB, pl.khi), None) for b0 in range(pl.klo, pl.khi, self.KB)]
    for bi, (b0, b1, only) in enumerate(blocks):
      # within a block: contraction index, then term, then row, so the right operand's row k
      # of every term is consumed right after it is formed
      es = [(e[1], t_idx, e[0], e) for t_idx, t in enumerate(pl.terms) for e in t.entries
            if (only is None or t_idx in only) and b0 <= e[1] < b1]
      for _, t_idx, _, e in sorted(es, key=lambda x: x[:3]):
        seq.append((bi, t_idx, e))
    self.blocks = blocks
    self.seq = seq
    # register stages: entries by previous-stage row, then row, after the streamed entries
    self.post_base = []
    for st_idx, st in enumerate(pl.post):
      U = st.U
      nd, kd = st.u_dims
      klo, khi = st.product.k_ranges[0]
      if U.is_sparse():
        items = U.packed.items()
      else:
        items = [(idx, U.offset(idx)) for idx in
                 ir.iterate_box([(U.origin[d], U.origin[d] + U.tensor.shape[d]) for d in range(2)])]
      st.entries = sorted((idx[kd] - st.prev_lo, idx[nd] - st.rows[0], pos) for idx, pos in items
                          if st.rows[0] <= idx[nd] < st.rows[1] and klo <= idx[kd] < khi)
      self.post_base.append(len(seq))
      for j, i, pos in st.entries:
        seq.append((None, len(pl.terms) + st_idx, (i, j, pos)))
    self.nu = len(seq)
    self.nu_pad = (self.nu + CHUNK - 1) // CHUNK * CHUNK

  def _loop_mode(self):
    """Dense shared operands with many entries run the contraction blocks as a runtime loop
    (the fully unrolled body would not fit the instruction cache)."""
    pl = self.pl
    if self.sched.stream_loop == 'never' or pl.post or pl.stacked:
      return False
    if any(t.U.is_sparse() for t in pl.terms):
      return False
    if any(tuple(t.product.k_ranges[0]) != (pl.klo, pl.khi) for t in pl.terms):
      return False
    if (pl.khi - pl.klo) % max(2, self.sched.stream_loop_k_block) != 0:
      return False
    if self.sched.stream_loop == 'always':
      return True
    return self.nu > 4096

  # ---------------------------------------------------------------- emission
  def emit(self) -> Emitted:
    pl = self.pl
    self._g0_declared = False
    self.order_entries()
    # streamed operands: one, or one per term when the terms stream different operands
    self.Ls, self.L_index = [], {}
    for L in ([t.L for t in pl.terms] if pl.stacked else [pl.L]):
      if L is not None and id(L.arg) not in self.L_index:
        self.L_index[id(L.arg)] = len(self.Ls)
        self.Ls.append(L)
    self.loop = self._loop_mode()
    if self.loop and self.KB != self.sched.stream_loop_k_block:
      # the runtime block loop keeps one block of every term in flight: smaller blocks
      self.KB = max(2, self.sched.stream_loop_k_block)
      self.order_entries()
      self.loop = self._loop_mode()
    # shared entries read two chunks per instruction, streamed rows two per instruction
    self.pair = bool(self.sched.stream_pair_reads) and not self.loop and self.real == 'double' \
        and not pl.stacked
    if self.pair:
      self.nu_pad = (self.nu + 2 * CHUNK - 1) // (2 * CHUNK) * (2 * CHUNK)
    self.rows_pad = (pl.rows[1] - pl.rows[0] + CHUNK - 1) // CHUNK * CHUNK
    if self.loop:
      # shared entries laid out column by column, each (column, term) padded to whole chunks
      self.nu_pad = (pl.khi - pl.klo) * len(pl.terms) * self.rows_pad
      self.u_source = self.sched.stream_u_source
      if self.u_source == 'auto':
        self.u_source = 'lds' if self.nu_pad * self.fp.size() <= 96 * 1024 else 'global'
      if self.u_source == 'global':
        self.nu_pad = 0
    w = self.w
    real = self.real
    W, G, q, KB = self.W, self.G, self.q, self.KB
    threads = WAVE * W
    # ring of the streamed operand: per wavefront 2 slots x pairs x (KB + 1) rows
    self.dma = False
    if pl.L is not None:
      self.l_q = pl.l_cols[1] - pl.l_cols[0]
      if pl.stacked and not self.own_columns():
        raise GenerationError('operands streamed one term after the other need the lanes to '
                              'load their own columns')
      self.pairs = G * self.l_q
      self.ps = KB + 2 if (self.pair and KB % 2 == 0) else KB + 1
      self.slot_words = self.pairs * self.ps
      ring = 2 * self.slot_words
      if self.allow_dma and self.dma_possible():
        # rows of the streamed operand go straight from memory to LDS: each load instruction
        # moves 16 bytes (rpl rows of its column) per lane, lane i's rows land at i * rpl of
        # the instruction's region; ring of NS slots of KB rows x 64 lanes
        self.dma = True
        self.rpl = 16 // self.fp.size()
        self.ns = max(2, min(self.sched.stream_dma_slots, len(self.blocks) + 1))
        self.ps = self.rpl
        self.slot_words = KB * WAVE
        ring = self.ns * self.slot_words
    else:
      ring = 0
    self.ring_words = ring
    lds_words = self.nu_pad + W * ring
    lds_bytes = lds_words * self.fp.size()
    # tables: position of every traversal entry in its U array, term in the top bits
    if not self.loop:
      tab = ', '.join(f'{(t_idx << 20) | e[2]}u' for (_, t_idx, e) in self.seq)
      # guarded: a grouped kernel repeats its members' tables in the same translation unit
      # named per row band: the bands of one kernel have different tables
      self.utab = f'{self.kname}_utab_r{pl.rows[0]}_{pl.rows[1]}'
      w(f'#ifndef TF_{self.utab}')
      w(f'#define TF_{self.utab}')
      w(f'static __device__ const unsigned {self.utab}[{max(1, self.nu)}] = '
        f'{{{tab or "0u"}}};')
      w('#endif')
    w('#ifndef TF_D2_TYPE')
    w('#define TF_D2_TYPE')
    w('typedef double tf_d2 __attribute__((ext_vector_type(2)));')
    w('#endif')
    params = ', '.join(kernel_params(self.prog, self.ctx))
    mw = max(1, self.sched.stream_min_waves)
    w(f'extern "C" __global__ void __launch_bounds__({threads}, {mw}) {self.kname}({params})')
    w('{')
    w._depth += 1
    w(f'// streaming schedule: {W} wavefront(s) per workgroup, {G} elements x {q} columns per '
      f'wavefront, contraction blocks of {KB}, {self.nu} shared entries in LDS')
    w(f'extern __shared__ __align__(16) {real} lds[];')
    w(f'{real}* __restrict__ ubuf = lds;')
    w(f'const int wave = threadIdx.x / {WAVE};')
    w(f'const int lane = threadIdx.x % {WAVE};')
    if ring:
      w(f'{real}* __restrict__ ring = lds + {self.nu_pad} + wave * {ring};')
    w(f'const int slot = lane / {q} < {G} ? lane / {q} : {G - 1};')
    w(f'const int col = lane / {q} < {G} ? lane - slot * {q} : 0;')
    # stage the shared entries
    if not self.loop:
      with w.block(f'for (int e = threadIdx.x; e < {self.nu}; e += {threads})'):
        w(f'const unsigned tv = {self.utab}[e];')
        w('const unsigned t = tv >> 20;')
        w('const unsigned pos = tv & 0xfffffu;')
        shared = [t.U for t in pl.terms] + [st.U for st in pl.post]
        expr = f'({real}){shared[-1].arg.name}[pos]'
        for t_idx in range(len(shared) - 2, -1, -1):
          expr = f'(t == {t_idx} ? ({real}){shared[t_idx].arg.name}[pos] : {expr})'
        if self.pair:
          # chunk pairs interleaved: entry j of chunks 2c and 2c + 1 side by side
          w(f'ubuf[((e >> 5) << 5) | ((e & 15) << 1) | ((e >> 4) & 1)] = {expr};')
        else:
          w(f'ubuf[e] = {expr};')
    elif self.u_source == 'lds':
      nt = len(pl.terms)
      rp = self.rows_pad
      nrows = pl.rows[1] - pl.rows[0]
      with w.block(f'for (int e = threadIdx.x; e < {self.nu_pad}; e += {threads})'):
        w(f'const int i = e % {rp};')
        w(f'const int t = (e / {rp}) % {nt};')
        w(f'const int k = e / {rp * nt} + {pl.klo};')
        expr = self.lit(0)
        for t_idx in range(nt - 1, -1, -1):
          U = pl.terms[t_idx].U
          nd, kd = pl.terms[t_idx].u_dims
          pos = (f'(i + {pl.rows[0] - U.origin[nd]}) * {U.strides[nd]} + (k - {U.origin[kd]}) * '
                 f'{U.strides[kd]}')
          expr = f'(t == {t_idx} ? ({real}){U.arg.name}[{pos}] : {expr})'
        w(f'ubuf[e] = i < {nrows} ? {expr} : {self.lit(0)};')
    w('__syncthreads();')
    self.per_lane_setup()
    w(f'const size_t ngroups = (numElements + {G - 1}) / {G};')
    w(f'const size_t stride = (size_t)gridDim.x * {W};')
    w('const size_t rounds = (ngroups + stride - 1) / stride;')
    w(f'const size_t efirst = ((size_t)blockIdx.x * {W} + wave) * {G};')
    w(f'const size_t estep = stride * {G};')
    if self.dma:
      self.dma_prologue()
    elif not self.loop:
      self.pipeline_prologue()
    with w.block('for (size_t r = 0; r < rounds; ++r)'):
      w('const size_t e0 = efirst + r * estep;')
      w('const size_t e0n = e0 + estep;')
      w('const size_t batchId = e0 + slot;')
      w(f'const bool active = (lane < {G * q}) && (batchId < numElements) && '
        '(flags == nullptr || flags[batchId] != 0u);')
      w('const size_t eSafe = active ? batchId : 0;')
      self.group_body()
    if self.dma:
      # loads issued for the group after the last one (clamped to element 0) must land first
      w('asm volatile("s_waitcnt vmcnt(0)" ::: "memory");')
    w._depth -= 1
    w('}')
    traffic = element_traffic(self.prog)
    sched = {'kind': 'stream', 'waves_per_block': W, 'k_block': KB, 'elements_per_wave': G,
             'columns': q, 'shared_entries': self.nu, 'lds_bytes': lds_bytes,
             'dpp': self.sched.dpp}
    if self.loop:
      sched['loop'] = True
    if self.pair:
      sched['pair'] = True
    if pl.L is not None and self.row_loads():
      sched['row_loads'] = True
    if self.prog.epilogue is not None:
      sched['epilogue'] = 'pointwise'
    if self.dma:
      sched['lds_dma_slots'] = self.ns
    fallback = None
    if self.dma:
      fallback = _Stream(self.prog, self.ctx, self.kname[len('kernel_'):], self.sched, self.pl,
                         allow_dma=False, suffix='_r').emit()
    return Emitted(kernel_name=self.kname, kernel_src=w.src(), block=(threads, 1, 1),
                   elements_per_block=G * W, lds_bytes=lds_bytes, emitter='tuned',
                   schedule=sched, bytes_per_element=traffic.per_element,
                   shared_bytes=traffic.shared,
                   grid_override='tf_grid',
                   extra_launch_setup=self.grid_setup(lds_bytes, threads),
                   fallback=fallback,
                   condition=self.dma_condition() if self.dma else '')

  def grid_setup(self, lds_bytes, threads):
    return persistent_grid(self.kname, lds_bytes, threads, self.G * self.W,
                           f'launcher_{self.base_name}')

  def _select(self, values):
    if all(v == values[0] for v in values):
      return str(values[0])
    expr = str(values[-1])
    for c in range(len(values) - 2, -1, -1):
      expr = f'(col == {c} ? {values[c]} : {expr})'
    return expr

  def per_lane_setup(self):
    """Per-lane slot tables of the fused temporaries (columns of L and packed positions of W)."""
    w = self.w
    pl = self.pl
    w('int lane15 = lane & 15;')
    self.slots = {}
    self.offsets = {}
    self.positions = {}
    for t_idx, t in enumerate(pl.terms):
      if t.kind != 'fused':
        continue
      prod = t.producer
      mlo, mhi = prod.k_ranges[0]
      c_lo = prod.n_ranges[1][0]
      rk, rn = t.w_dims
      cols = []
      for c in range(self.q):
        ms = []
        for m in range(mlo, mhi):
          idx = [0, 0]
          idx[rk], idx[rn] = m, c_lo + c
          if t.W.is_sparse():
            pos = t.W.packed.get(tuple(idx))
          else:
            pos = t.W.offset(idx)
          if pos is not None:
            ms.append((m, pos))
        cols.append(ms)
      nslots = max(len(c) for c in cols)
      entries = []
      for s in range(nslots):
        ms = tuple(c[s][0] if s < len(c) else (c[0][0] if c else mlo) for c in cols)
        ps = [c[s][1] if s < len(c) else -1 for c in cols]
        v = self.fresh('sl')
        self.positions[v] = ps
        # ring offset of the slot's column within the lane's element: pair * (KB + 1); terms
        # whose patterns pick the same columns share the offset, and so the LDS reads
        if ms not in self.offsets:
          o = self.fresh('off')
          w(f'int {o} = (slot * {self.l_q} + ({self._select(list(ms))} - {self.pl.l_cols[0]})) * '
            f'{self.ps};')
          self.offsets[ms] = o
        entries.append((v, self.offsets[ms]))
      self.slots[t_idx] = entries

  def group_body(self):
    w = self.w
    pl = self.pl
    real = self.real
    rows = pl.rows[1] - pl.rows[0]
    # per-element pointers
    for a in self.prog.args:
      if a.kind != ir.SRC_GLOBAL or not a.per_element:
        continue
      ty = f'{"const " if a.const else ""}{a.dtype}'
      if a.addressing == Addressing.STRIDED:
        w(f'{ty}* __restrict__ g_{a.name} = {a.name} + eSafe * {a.element_stride}ull + '
          f'{a.name}_extraOffset;')
      else:
        w(f'{ty}* __restrict__ g_{a.name} = {a.name}[eSafe] + {a.name}_extraOffset;')
    # weights of the fused temporaries; their positions are recomputed from an opaque copy of
    # the lane's column each group, so they do not occupy registers across the loop
    if self.slots:
      w('int colr = col;')
      w('asm volatile("" : "+v"(colr));')
    for t_idx, slots in self.slots.items():
      t = pl.terms[t_idx]
      for v, _ in slots:
        w(f'const int {v}_p = {self._select(self.positions[v]).replace("col ==", "colr ==")};')
        w(f'const {real} {v}_x = ({real})g_{t.W.arg.name}[{v}_p < 0 ? 0 : {v}_p];')
        w(f'const {real} {v}_v = {v}_p >= 0 ? {v}_x : {self.lit(0)};')
    # destination column
    acc = 'acc'
    w(f'{real} {acc}[{rows}];')
    dp = pl.dest_product
    o0, o1 = dp.dest_origin
    s0, s1 = dp.dest_strides
    for i in range(rows):
      off = f'{(pl.rows[0] + i - o0) * s0} + (col + {pl.cols[0] - o1}) * {s1}'
      if pl.init == 'load' and not pl.post:
        old = f'({real})g_{pl.dest.name}[{off}]'
        w(f'{acc}[{i}] = {old if pl.beta == 1.0 else f"{self.lit(pl.beta)} * {old}"};')
      else:
        w(f'{acc}[{i}] = {self.lit(0)};')
    if self.loop:
      self.loop_blocks(acc)
      with w.block('if (active)'):
        self.epilogue(acc, pl.rows[0], rows)
        for i in range(rows):
          off = f'{(pl.rows[0] + i - o0) * s0} + (col + {pl.cols[0] - o1}) * {s1}'
          w(f'g_{pl.dest.name}[{off}] = {acc}[{i}];')
      return
    nb = len(self.blocks)
    if self.dma:
      self.dma_blocks(acc)
      with w.block('if (active)'):
        self.epilogue(acc, pl.rows[0], rows)
        for i in range(rows):
          off = f'{(pl.rows[0] + i - o0) * s0} + (col + {pl.cols[0] - o1}) * {s1}'
          w(f'g_{pl.dest.name}[{off}] = {acc}[{i}];')
      return
    D = self.D
    self._nxt_ptr_done = False
    self._nxt_ptrs = set()
    if self.own_columns():
      self.declare_lane_pointers('e0', 'cur')
    if self.row_loads():
      self.row_offsets('e0', 'ro')
    for bi, (b0, b1, only) in enumerate(self.blocks):
      w(f'// block [{b0}, {b1})' + (f' of term {only[0]}' if only else ''))
      # issue the loads of the block D ahead (this group's, or the next group's)
      tb = bi + D
      if tb < nb:
        self.pipeline_load(tb, 'e0', D - 1)
      else:
        self.pipeline_load(tb - nb, 'e0n', D - 1)
      self.global_rows = {bi: self.pipe_global[0]}
      if self.sched.stream_lds_lookahead > 0:
        self.compute_block_pipelined(bi, b0, acc)
      else:
        self.compute_block(bi, b0, acc)
      # block bi + 1 (or the next group's first block) into the ring
      nxt = bi + 1 if bi + 1 < nb else 0
      self.pipeline_write(nxt)
      self.pipeline_shift()
    self.pipeline_realign()
    if pl.post:
      self.post_stages(acc)
      return
    with w.block('if (active)'):
      self.epilogue(acc, pl.rows[0], rows)
      for i in range(rows):
        off = f'{(pl.rows[0] + i - o0) * s0} + (col + {pl.cols[0] - o1}) * {s1}'
        w(f'g_{pl.dest.name}[{off}] = {acc}[{i}];')

  def post_stages(self, acc):
    """Register stages: stage s accumulates U_s (LDS chunks, DPP broadcast) times the previous
    stage's column; the last one starts from the destination when it accumulates and is
    stored."""
    w = self.w
    pl = self.pl
    real = self.real
    prev = acc
    ahead = max(1, self.sched.stream_lds_lookahead)
    for s_idx, st in enumerate(pl.post):
      last = s_idx == len(pl.post) - 1
      name = f'acp{s_idx}'
      nrows = st.rows[1] - st.rows[0]
      dp = st.product
      o0, o1 = dp.dest_origin
      s0, s1 = dp.dest_strides
      w(f'{real} {name}[{nrows}];')
      for i in range(nrows):
        if last and pl.final_init == 'load':
          off = f'{(st.rows[0] + i - o0) * s0} + (col + {pl.cols[0] - o1}) * {s1}'
          old = f'({real})g_{pl.dest.name}[{off}]'
          w(f'{name}[{i}] = {old if pl.final_beta == 1.0 else f"{self.lit(pl.final_beta)} * {old}"};')
        else:
          w(f'{name}[{i}] = {self.lit(0)};')
      base = self.post_base[s_idx]
      order = []
      for n in range(len(st.entries)):
        c = (base + n) // CHUNK
        if c not in order:
          order.append(c)
      cvar = {}
      issued = [0]

      def issue(upto):
        while issued[0] < min(upto, len(order)):
          c = order[issued[0]]
          if c not in cvar:
            if self.pair:
              base = c - c % 2
              v = self.fresh('pp')
              w(f'const tf_d2 {v} = {self.chunk_pair_ref(base)};')
              cvar[base], cvar[base + 1] = f'{v}.x', f'{v}.y'
            else:
              v = self.fresh('pc')
              w(f'const {real} {v} = {self.chunk_ref(c)};')
              cvar[c] = v
          issued[0] += 1

      w('asm volatile("" : "+v"(lane15));')
      issue(ahead)
      cur = None
      for n, (j, i, pos) in enumerate(st.entries):
        c, lane_in = divmod(base + n, CHUNK)
        if c != cur:
          if cur is not None:
            w('__builtin_amdgcn_sched_barrier(0);')
          issue(order.index(c) + 1 + ahead)
          cur = c
        self.fma(f'{name}[{i}]', cvar[c], lane_in, f'{prev}[{j}]')
      w('__builtin_amdgcn_sched_barrier(0);')
      prev = name
    st = pl.post[-1]
    dp = st.product
    o0, o1 = dp.dest_origin
    s0, s1 = dp.dest_strides
    with w.block('if (active)'):
      self.epilogue(prev, st.rows[0], st.rows[1] - st.rows[0])
      for i in range(st.rows[1] - st.rows[0]):
        off = f'{(st.rows[0] + i - o0) * s0} + (col + {pl.cols[0] - o1}) * {s1}'
        w(f'g_{pl.dest.name}[{off}] = {prev}[{i}];')

  def epilogue(self, acc, row_lo, nrows):
    """Pointwise statements fused into the chain (``program.epilogue``): run on the lane's
    destination column in registers, row by row, before it is stored (active lanes only)."""
    pw = self.prog.epilogue
    if pw is None:
      return
    from . import pointwise as pwm
    pwm.check_epilogue(pw, self.pl.dest.tensor, (row_lo, row_lo + nrows),
                       (self.pl.cols[0], self.pl.cols[0] + self.q))
    w = self.w
    dest = self.pl.dest.tensor
    for i in range(nrows):
      sfx = f'_r{i}'
      with w.block(''):
        for c in pw.cells:
          if c.tensor is dest:
            w(f'{c.dtype} c{c.index}{sfx} = {acc}[{i}];')
            w(f'bool w{c.index}{sfx} = false;')
            w(f'(void)w{c.index}{sfx};')
          else:
            # another argument, read at this point: row i of the window, the lane's column
            off = pwm.epilogue_address(c, [str(i), 'col'])
            base = f'g_{c.arg.name}' if c.arg.per_element else c.arg.name
            w(f'const {c.dtype} c{c.index}{sfx} = ({c.dtype}){base}[{off}];')
        for t in pw.temps:
          dt = pw.fp_type if t.datatype is None else t.datatype
          w(f'{dt} {t.name}{sfx} = {dt.literal(0)};')
        pwm._Code(pw, w, sfx).statements(pw.statements)
        for c in pw.cells:
          if c.tensor is dest:
            w(f'{acc}[{i}] = c{c.index}{sfx};')

  # ---- cross-group prefetch pipeline (unrolled blocks) ----
  def _pipe_steps(self, nr):
    return (self.pairs * nr + WAVE - 1) // WAVE

  def pipeline_prologue(self):
    """Declare the pipeline registers and fill them with the first group's blocks 0..D-1;
    block 0 goes to the ring."""
    w = self.w
    pl = self.pl
    self.D = max(1, min(self.sched.stream_prefetch, len(self.blocks)))
    D = self.D
    nrmax = self.KB
    self.pipe_names = []
    for j in range(D):
      if pl.L is not None:
        size = nrmax if self.own_columns() else self._pipe_steps(nrmax)
        w(f'{self.real} pf{j}[{size}];')
    if self.row_loads():
      self.row_offsets('efirst', 'ro')
      self.pipe_names.append(f'pf{j}')
    if self.own_columns():
      w(f'const int lane_ro = (slot * {self.l_q} + col) * {self.ps};')
    self.gterms = [ti for ti, t in enumerate(pl.terms) if t.kind == 'global']
    for j in range(D):
      for ti in self.gterms:
        w(f'{self.real} pg{j}_{ti}[{nrmax}];')
    # rotation state: python-side mapping of logical pipeline position -> register set name
    self.pipe_rot = list(range(D))
    self.pipe_block = [None] * D      # block index held by each logical position
    w('{')
    w._depth += 1
    w('const size_t e0 = efirst;')
    w('const size_t e0n = efirst + estep;')
    self._nxt_ptrs = set()
    if self.own_columns():
      self.declare_lane_pointers('e0', 'cur')
    for j in range(D):
      if j < len(self.blocks):
        self.pipeline_load(j, 'e0', j, declare=False)
    w._depth -= 1
    w('}')
    self.pipeline_write(0)
    self.pipeline_shift()
    self.pipe_rot_entry = list(self.pipe_rot)

  # ---- direct memory-to-LDS ring (rows of the streamed operand bypass registers) ----
  def dma_possible(self):
    """Direct loads need 16-byte aligned row groups: rows contiguous, column and element
    strides multiples of 16 bytes, blocks made of whole row groups (the base pointer's
    alignment is checked by the launcher)."""
    pl = self.pl
    L = pl.L
    if not (self.sched.stream_dma and not self.loop and self.own_columns()) or self.pl.stacked:
      return False
    if any(t.kind == 'global' for t in pl.terms) or len(self.blocks) < 2:
      return False
    size = self.fp.size()
    rpl = 16 // size
    return (L.strides[0] == 1 and (L.strides[1] * size) % 16 == 0 and
            (L.arg.element_stride * size) % 16 == 0 and self.KB % rpl == 0 and
            (pl.khi - pl.klo) % rpl == 0 and (pl.klo - L.origin[0]) % rpl == 0 and
            ((pl.l_cols[0] - L.origin[1]) * L.strides[1] * size) % 16 == 0)

  def dma_condition(self):
    a = self.pl.L.arg
    return f'(reinterpret_cast<uintptr_t>({a.name} + {a.name}_extraOffset) % 16u == 0u)'

  def chunk_ref(self, c):
    """LDS expression of lane15's entry of shared-operand chunk c."""
    if self.pair:
      return f'ubuf[{(c // 2) * 2 * CHUNK + c % 2} + lane15 * 2]'
    return f'ubuf[{c * CHUNK} + lane15]'

  def chunk_pair_ref(self, c):
    """16-byte LDS expression of chunks (c, c + 1) for an even c (pair layout)."""
    return f'*reinterpret_cast<const tf_d2*>(&ubuf[{(c // 2) * 2 * CHUNK} + lane15 * 2])'

  def ring_ref(self, ring_slot, o, kk):
    if self.dma:
      r = self.rpl
      return f'rb[{o} + {(kk // r) * WAVE * r + kk % r}]'
    return f'ring[{ring_slot * self.slot_words} + {o} + {kk}]'

  def dma_issue(self, bi, ptr, slot_expr):
    """Loads of block ``bi``'s rows of the lane's column (pointer ``ptr``) into ring slot
    ``slot_expr``: one instruction per row, lane i's value lands at row base + i."""
    w = self.w
    L = self.pl.L
    b0, b1 = self._block_rows(bi)
    r = self.rpl
    for j in range((b1 - b0) // r):
      goff = b0 + j * r - L.origin[0]
      w(f'__builtin_amdgcn_global_load_lds((const void*)({ptr} + {goff}), '
        f'(tf_lds_t*)(ring + ({slot_expr}) * {self.slot_words} + {j * WAVE * r}), 16, 0, 0);')

  def dma_prologue(self):
    w = self.w
    nb = len(self.blocks)
    self.D = self.ns - 1
    w('typedef __attribute__((address_space(3))) void tf_lds_t;')
    with w.block(''):
      w(f'const {self.pl.L.arg.dtype}* __restrict__ lp0 = {self.lane_pointer("efirst")};')
      for j in range(self.D):
        self.dma_issue(j, 'lp0', str(j))

  def dma_blocks(self, acc):
    """Blocks of one element group: wait for block bi's rows, start the loads of the block
    D ahead (this group's or the next one's) into the slot freed by block bi - 1, compute."""
    w = self.w
    pl = self.pl
    nb = len(self.blocks)
    D, NS, KB = self.D, self.ns, self.KB
    L = pl.L
    w(f'const {L.arg.dtype}* __restrict__ lp_cur = {self.lane_pointer("e0")};')
    w(f'const {L.arg.dtype}* __restrict__ lp_nxt = {self.lane_pointer("e0n")};')
    w(f'const unsigned gs = (unsigned)((r * {nb}) % {NS});')
    # at most the loads of the D - 1 younger blocks may still be in flight
    pending = min(63, (D - 1) * (KB // self.rpl))
    for bi, (b0, b1, _) in enumerate(self.blocks):
      w(f'// block [{b0}, {b1})')
      w(f'asm volatile("s_waitcnt vmcnt({pending})" ::: "memory");')
      w('{')
      w._depth += 1
      w(f'{self.real}* __restrict__ rb = ring + ((gs + {bi}u) % {NS}u) * {self.slot_words};')
      tb = bi + D
      if tb < nb:
        self.dma_issue(tb, 'lp_cur', f'(gs + {tb}u) % {NS}u')
      else:
        self.dma_issue(tb - nb, 'lp_nxt', f'(gs + {tb}u) % {NS}u')
      if self.sched.stream_lds_lookahead > 0:
        self.compute_block_pipelined(bi, b0, acc)
      else:
        self.compute_block(bi, b0, acc)
      w._depth -= 1
      w('}')
      w('__builtin_amdgcn_sched_barrier(0);')

  def _set(self, pos):
    return self.pipe_rot[pos]

  def own_columns(self):
    """The streamed operand has as many columns as the output: lane (slot, col) loads the
    rows of its own element's column, so addresses are one per-lane pointer plus constants."""
    return self.pl.L is not None and self.l_q == self.q and \
        self.pl.L.arg.addressing == Addressing.STRIDED and self.sched.stream_own_columns and \
        not self.row_loads()

  def row_loads(self):
    """Consecutive lanes load consecutive rows of one column (KB rows x 8 columns per load
    instruction for doubles), from a wavefront-uniform base and per-lane offsets computed once
    per element group. Needs full blocks."""
    pl = self.pl
    return (pl.L is not None and pl.L.arg.addressing == Addressing.STRIDED and
            self.sched.stream_row_loads and not self.loop and not pl.stacked and
            (pl.khi - pl.klo) % self.KB == 0 and
            not any(t.kind == 'global' for t in pl.terms))

  def row_offsets(self, ebase, name):
    """Declare ``{name}_q`` (wavefront-uniform base of the element group at ``ebase``, clamped
    into the batch) and per-step lane offsets ``{name}{s}`` of the rows the lane loads."""
    w = self.w
    pl = self.pl
    L = pl.L
    a = L.arg
    KB = self.KB
    total = self.pairs * KB
    w(f'const size_t {name}_e = {ebase} < numElements ? {ebase} : numElements - 1;')
    w(f'const {a.dtype}* __restrict__ {name}_q = {a.name} + {a.name}_extraOffset + '
      f'{name}_e * {a.element_stride}ull;')
    for s_ in range(self._pipe_steps(KB)):
      idx = f'(lane + {s_ * WAVE})' if s_ * WAVE + WAVE <= total else \
          f'(lane + {s_ * WAVE} < {total} ? lane + {s_ * WAVE} : {total - 1})'
      w(f'int {name}{s_};')
      with w.block(''):
        w(f'const int ix = {idx};')
        w(f'const int pr = ix / {KB};')
        w(f'const int el = pr / {self.l_q};')
        w(f'const int cc = pr - el * {self.l_q};')
        w(f'const int ev = ({name}_e + el < numElements ? el : (int)(numElements - 1 - {name}_e));')
        w(f'{name}{s_} = ev * {a.element_stride} + (cc + {pl.l_cols[0] - L.origin[1]}) * '
          f'{L.strides[1]} + (ix - pr * {KB} + {pl.klo - L.origin[0]}) * {L.strides[0]};')

  def lane_pointer(self, ebase, L=None):
    """Expression of the lane's pointer into its element's column of L for group ``ebase``."""
    L = L or self.pl.L
    a = L.arg
    e = f'({ebase} + slot < numElements ? {ebase} + slot : 0)'
    return (f'({a.name} + {e} * {a.element_stride}ull + {a.name}_extraOffset + '
            f'(col + {self.pl.l_cols[0] - L.origin[1]}) * {L.strides[1]})')

  def lp_name(self, which, li):
    """Name of the lane's pointer into streamed operand ``li`` ('cur' or 'nxt' group)."""
    return f'lp_{which}' + (str(li) if li else '')

  def declare_lane_pointers(self, ebase, which):
    for li, L in enumerate(self.Ls):
      w_ = self.w
      w_(f'const {L.arg.dtype}* __restrict__ {self.lp_name(which, li)} = '
         f'{self.lane_pointer(ebase, L)};')

  def pipeline_load(self, bi, ebase, pos, declare=False):
    """Registers of logical pipeline position ``pos`` <- block ``bi`` of the group at
    ``ebase`` (streamed operand rows and per-element right-operand rows)."""
    w = self.w
    pl = self.pl
    reg = self._set(pos)
    b0, b1 = self._block_rows(bi)
    nr = b1 - b0
    if pl.L is not None and self.own_columns():
      li, L = self._block_L(bi)
      only = self.blocks[bi][2]
      shift = pl.terms[only[0]].l_shift if only else 0
      ptr = self.lp_name('cur' if ebase == 'e0' else 'nxt', li)
      if ebase != 'e0' and li not in self._nxt_ptrs:
        w(f'const {L.arg.dtype}* __restrict__ {ptr} = {self.lane_pointer("e0n", L)};')
        self._nxt_ptrs.add(li)
      for kk in range(nr):
        w(f'pf{reg}[{kk}] = ({self.real}){ptr}[{(b0 + kk + shift - L.origin[0]) * L.strides[0]}];')
    elif pl.L is not None and self.row_loads():
      L = pl.L
      if ebase == 'e0':
        name = 'ro'
      else:
        if not getattr(self, '_nxt_ptr_done', False):
          self.row_offsets('e0n', 'rn')
          self._nxt_ptr_done = True
        name = 'rn'
      for s_ in range(self._pipe_steps(nr)):
        w(f'pf{reg}[{s_}] = ({self.real}){name}_q[{name}{s_} + {(b0 - pl.klo) * L.strides[0]}];')
    elif pl.L is not None:
      L = pl.L
      total = self.pairs * nr
      vol = L.arg.element_stride
      for s_ in range(self._pipe_steps(nr)):
        if s_ * WAVE + WAVE - 1 < total:
          idx = f'lane + {s_ * WAVE}'
        else:
          idx = f'(lane + {s_ * WAVE} < {total} ? lane + {s_ * WAVE} : {total - 1})'
        pr = f'(({idx}) / {nr})'
        kk = f'(({idx}) % {nr})'
        el = f'({pr} / {self.l_q})'
        cc = f'({pr} % {self.l_q})'
        e = f'({ebase} + {el} < numElements ? {ebase} + {el} : 0)'
        if L.arg.addressing == Addressing.STRIDED:
          base = f'{L.arg.name} + {e} * {vol}ull + {L.arg.name}_extraOffset'
        else:
          base = f'{L.arg.name}[{e}] + {L.arg.name}_extraOffset'
        goff = (f'({b0 - L.origin[0]} + {kk}) * {L.strides[0]} + '
                f'({pl.l_cols[0] - L.origin[1]} + {cc}) * {L.strides[1]}')
        w(f'pf{reg}[{s_}] = ({self.real})({base})[{goff}];')
    for ti in self.gterms:
      t = pl.terms[ti]
      B = t.B
      kd, nd = t.b_dims
      a = B.arg
      eb = f'({ebase} + slot < numElements ? {ebase} + slot : 0)'
      if a.addressing == Addressing.STRIDED:
        base = f'{a.name} + {eb} * {a.element_stride}ull + {a.name}_extraOffset'
      else:
        base = f'{a.name}[{eb}] + {a.name}_extraOffset'
      for j, k in enumerate(range(b0, b1)):
        off = (f'({k - B.origin[kd]}) * {B.strides[kd]} + (col + {pl.cols[0] - B.origin[nd]}) '
               f'* {B.strides[nd]}')
        w(f'pg{reg}_{ti}[{j}] = ({self.real})({base})[{off}];')
    self.pipe_block[pos] = bi

  def pipeline_write(self, bi):
    """The oldest pipeline position holds block ``bi``: put its streamed rows in the ring and
    expose its right-operand rows to the compute of that block."""
    w = self.w
    pl = self.pl
    reg = self._set(0)
    b0, b1 = self._block_rows(bi)
    nr = b1 - b0
    if pl.L is not None and self.own_columns():
      slot_ = bi % 2
      for kk in range(nr):
        w(f'ring[{slot_ * self.slot_words} + lane_ro + {kk}] = pf{reg}[{kk}];')
      w('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");')
      w('__builtin_amdgcn_wave_barrier();')
      w('__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')
    elif pl.L is not None:
      total = self.pairs * nr
      slot_ = bi % 2
      for s_ in range(self._pipe_steps(nr)):
        if s_ * WAVE + WAVE - 1 < total:
          idx = f'lane + {s_ * WAVE}'
        else:
          idx = f'(lane + {s_ * WAVE} < {total} ? lane + {s_ * WAVE} : {total - 1})'
        w(f'ring[{slot_ * self.slot_words} + (({idx}) / {nr}) * {self.ps} + (({idx}) % {nr})] = '
          f'pf{reg}[{s_}];')
      w('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");')
      w('__builtin_amdgcn_wave_barrier();')
      w('__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')
    # right-operand rows of block bi, as plain values for compute_block. Block 0 is written
    # before the element-group loop and again at the end of every group for the next one, and
    # its compute runs at the top of the loop body: its values live in variables carried across
    # iterations (declared by the first write, assigned by the later ones).
    vals = {}
    for ti in self.gterms:
      for j, k in enumerate(range(b0, b1)):
        if bi == 0:
          v = f'g0r_{ti}_{j}'
          if self._g0_declared:
            w(f'{v} = pg{reg}_{ti}[{j}];')
          else:
            w(f'{self.real} {v} = pg{reg}_{ti}[{j}];')
        else:
          v = self.fresh('gr')
          w(f'const {self.real} {v} = pg{reg}_{ti}[{j}];')
        vals[(ti, k)] = v
    if bi == 0:
      self._g0_declared = True
    self.pipe_global = [vals]

  def pipeline_realign(self):
    """Back-edge of the element-group loop: the pending blocks must sit in the register sets
    the loop entry expects. When the block count is not a multiple of the prefetch distance
    the rotation has drifted, and the pending sets are copied back."""
    w = self.w
    entry = self.pipe_rot_entry
    now = self.pipe_rot
    moves = [(now[p], entry[p]) for p in range(self.D - 1) if now[p] != entry[p]]
    if not moves:
      return
    steps = self._pipe_steps(self.KB) if self.pl.L is not None else 0
    w('{')
    w._depth += 1
    for src, dst in moves:
      if self.pl.L is not None:
        w(f'{self.real} tpf{src}[{steps}];')
        for s_ in range(steps):
          w(f'tpf{src}[{s_}] = pf{src}[{s_}];')
      for ti in self.gterms:
        w(f'{self.real} tpg{src}_{ti}[{self.KB}];')
        for j in range(self.KB):
          w(f'tpg{src}_{ti}[{j}] = pg{src}_{ti}[{j}];')
    for src, dst in moves:
      if self.pl.L is not None:
        for s_ in range(steps):
          w(f'pf{dst}[{s_}] = tpf{src}[{s_}];')
      for ti in self.gterms:
        for j in range(self.KB):
          w(f'pg{dst}_{ti}[{j}] = tpg{src}_{ti}[{j}];')
    w._depth -= 1
    w('}')
    self.pipe_rot = list(entry)

  def pipeline_shift(self):
    """Logical positions move down by one; the freed register set becomes the newest."""
    self.pipe_rot = self.pipe_rot[1:] + self.pipe_rot[:1]
    self.pipe_block = self.pipe_block[1:] + [None]

  # ---- runtime block loop (dense shared operands) ----
  def loop_blocks(self, acc):
    """The contraction as a runtime loop over blocks of KB rows. The streamed operand's next
    block and the next block's rows of per-element right operands are loaded one iteration
    ahead; the ring slot alternates with the block parity."""
    w = self.w
    pl = self.pl
    real = self.real
    KB = self.KB
    nb = (pl.khi - pl.klo) // KB
    nrows = pl.rows[1] - pl.rows[0]
    nt = len(pl.terms)
    fused_terms = [ti for ti, t in enumerate(pl.terms) if t.kind == 'fused']
    offs = sorted(set(o for ti in fused_terms for _, o in self.slots[ti]))
    if pl.L is not None:
      self.loop_load(0, 'pf')
      self.loop_write('pf', '0')
    gterms = [ti for ti, t in enumerate(pl.terms) if t.kind == 'global']
    for ti in gterms:
      for j in range(KB):
        w(f'{real} gb{ti}_{j} = {self.global_row_expr(ti, f"{pl.klo + j}")};')
    w('#pragma unroll 1')
    with w.block(f'for (int kb = 0; kb < {nb}; ++kb)'):
      w(f'const int k0 = {pl.klo} + kb * {KB};')
      w(f'const int rs = (kb & 1) * {self.slot_words if pl.L is not None else 0};')
      if pl.L is not None:
        self.loop_load(f'(kb + 1 < {nb} ? kb + 1 : kb)', 'nx')
      for ti in gterms:
        for j in range(KB):
          w(f'const {real} cur{ti}_{j} = gb{ti}_{j};')
        for j in range(KB):
          w(f'gb{ti}_{j} = {self.global_row_expr(ti, f"(kb + 1 < {nb} ? k0 + {KB + j} : k0 + {j})")};')
      w(f'asm volatile("" : {", ".join(f"{chr(34)}+v{chr(34)}({o})" for o in ["lane15"] + offs)});')
      for j in range(KB):
        rv = {}
        if fused_terms:
          qv = {}
          for o in offs:
            v = self.fresh('q')
            w(f'const {real} {v} = ring[rs + {o} + {j}];')
            qv[o] = v
        for ti in range(nt):
          t = pl.terms[ti]
          if t.kind == 'fused':
            slots = self.slots[ti]
            v = self.fresh('r')
            w(f'{real} {v} = {qv[slots[0][1]]} * {slots[0][0]}_v;')
            for sv, o in slots[1:]:
              w(f'{v} = __builtin_fma({qv[o]}, {sv}_v, {v});')
            rv[ti] = v
          else:
            rv[ti] = f'cur{ti}_{j}'
        for ti in range(nt):
          for c in range(self.rows_pad // CHUNK):
            cv = self.fresh('ch')
            w(f'const {real} {cv} = {self.loop_chunk(ti, f"k0 + {j}", c)};')
            for jj in range(CHUNK):
              i = c * CHUNK + jj
              if i >= nrows:
                break
              self.fma(f'{acc}[{i}]', cv, jj, rv[ti])
      if pl.L is not None:
        self.loop_write('nx', '((kb + 1) & 1)')

  def loop_chunk(self, ti, kexpr, c):
    pl = self.pl
    t = pl.terms[ti]
    if self.u_source == 'lds':
      nt = len(pl.terms)
      return (f'ubuf[((({kexpr}) - {pl.klo}) * {nt} + {ti}) * {self.rows_pad} + {c * CHUNK} + '
              f'lane15]')
    U = t.U
    nd, kd = t.u_dims
    nrows = pl.rows[1] - pl.rows[0]
    last = nrows - 1
    # rows past the window repeat the last row (loaded, never used)
    row = f'({c * CHUNK} + lane15 < {nrows} ? {c * CHUNK} + lane15 : {last})'
    return (f'({self.real}){U.arg.name}[({row} + {pl.rows[0] - U.origin[nd]}) * {U.strides[nd]} + '
            f'(({kexpr}) - {U.origin[kd]}) * {U.strides[kd]}]')

  def global_row_expr(self, ti, kexpr):
    pl = self.pl
    B = pl.terms[ti].B
    kd, nd = pl.terms[ti].b_dims
    return (f'({self.real})g_{B.arg.name}[(({kexpr}) - {B.origin[kd]}) * {B.strides[kd]} + '
            f'(col + {pl.cols[0] - B.origin[nd]}) * {B.strides[nd]}]')

  def loop_load(self, kbexpr, tag):
    """Registers ``<tag>_<s>`` <- rows [kb*KB, kb*KB+KB) of L for the wavefront's pairs."""
    w = self.w
    pl = self.pl
    L = pl.L
    KB = self.KB
    total = self.pairs * KB
    steps = (total + WAVE - 1) // WAVE
    vol = L.arg.element_stride
    self.loop_names = getattr(self, 'loop_names', {})
    names = []
    for s_ in range(steps):
      v = f'{tag}_{s_}'
      if s_ * WAVE + WAVE - 1 < total:
        idx = f'lane + {s_ * WAVE}'
      else:
        idx = f'(lane + {s_ * WAVE} < {total} ? lane + {s_ * WAVE} : {total - 1})'
      pr = f'(({idx}) / {KB})'
      kk = f'(({idx}) % {KB})'
      el = f'({pr} / {self.l_q})'
      cc = f'({pr} % {self.l_q})'
      e = f'(e0 + {el} < numElements ? e0 + {el} : 0)'
      if L.arg.addressing == Addressing.STRIDED:
        base = f'{L.arg.name} + {e} * {vol}ull + {L.arg.name}_extraOffset'
      else:
        base = f'{L.arg.name}[{e}] + {L.arg.name}_extraOffset'
      goff = (f'({pl.klo - L.origin[0]} + ({kbexpr}) * {KB} + {kk}) * {L.strides[0]} + '
              f'({pl.l_cols[0] - L.origin[1]} + {cc}) * {L.strides[1]}')
      if tag not in self.loop_names:
        w(f'{self.real} {v};')
      w(f'{v} = ({self.real})({base})[{goff}];')
      names.append((v, pr, kk))
    self.loop_names[tag] = names

  def loop_write(self, tag, slot_expr):
    w = self.w
    for v, pr, kk in self.loop_names[tag]:
      w(f'ring[({slot_expr}) * {self.slot_words} + {pr} * {self.ps} + {kk}] = {v};')
    w('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");')
    w('__builtin_amdgcn_wave_barrier();')
    w('__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')

  # ---- ring of the streamed operand ----
  def _block_rows(self, bi):
    b0, b1, _ = self.blocks[bi]
    return b0, b1

  def _block_L(self, bi):
    """(index into self.Ls, streamed operand) of block bi's rows."""
    only = self.blocks[bi][2]
    if not self.pl.stacked:
      return 0, self.pl.L
    L = self.pl.terms[only[0]].L
    return self.L_index[id(L.arg)], L

  def load_block(self, bi, stage):
    """Load the block's rows of L for every (element, column) pair of this wavefront into
    registers (one value per lane per step); ``stage`` writes them to the ring right away."""
    w = self.w
    pl = self.pl
    L = pl.L
    b0, b1 = self._block_rows(bi)
    nr = b1 - b0
    total = self.pairs * nr
    steps = (total + WAVE - 1) // WAVE
    names = []
    vol = L.arg.element_stride
    for s in range(steps):
      v = self.fresh('rg')
      # flat index -> (pair, row): rows fastest so a lane's neighbours read neighbouring rows;
      # lanes past the end repeat the last entry (same value to the same LDS word)
      last = min(total - 1, s * WAVE + WAVE - 1)
      if s * WAVE + WAVE - 1 < total:
        w(f'const int {v}_i = lane + {s * WAVE};')
      else:
        w(f'const int {v}_i = lane + {s * WAVE} < {total} ? lane + {s * WAVE} : {total - 1};')
      w(f'const int {v}_pr = {v}_i / {nr};')
      w(f'const int {v}_k = {v}_i - {v}_pr * {nr};')
      w(f'const int {v}_el = {v}_pr / {self.l_q};')
      w(f'const int {v}_c = {v}_pr - {v}_el * {self.l_q};')
      w(f'const size_t {v}_e = e0 + {v}_el < numElements ? e0 + {v}_el : 0;')
      if L.arg.addressing == Addressing.STRIDED:
        base = f'{L.arg.name} + {v}_e * {vol}ull + {L.arg.name}_extraOffset'
      else:
        base = f'{L.arg.name}[{v}_e] + {L.arg.name}_extraOffset'
      goff = (f'({b0 - L.origin[0]} + {v}_k) * {L.strides[0]} + '
              f'({pl.l_cols[0] - L.origin[1]} + {v}_c) * {L.strides[1]}')
      w(f'const {self.real} {v} = ({self.real})({base})[{goff}];')
      names.append(v)
    self.pending = (bi, names)
    if stage:
      self.write_block(bi)

  def write_block(self, bi):
    w = self.w
    pbi, names = self.pending
    assert pbi == bi
    b0, b1 = self._block_rows(bi)
    nr = b1 - b0
    ring_slot = bi % 2
    for v in names:
      w(f'ring[{ring_slot * self.slot_words} + {v}_pr * {self.ps} + {v}_k] = {v};')
    # LDS accesses of one wavefront are performed in order; the fence keeps the compiler from
    # moving the block's reads above these writes
    w('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");')
    w('__builtin_amdgcn_wave_barrier();')
    w('__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')

  def load_global_rows(self, bi):
    """Rows of the block of every per-element right operand, loaded one block ahead."""
    w = self.w
    pl = self.pl
    b0, b1 = self._block_rows(bi)
    out = {}
    for t_idx, t in enumerate(pl.terms):
      if t.kind != 'global':
        continue
      B = t.B
      kd, nd = t.b_dims
      for k in range(b0, b1):
        if not any(e[1] == k for e in t.entries):
          continue
        v = self.fresh('gr')
        off = (f'({k - B.origin[kd]}) * {B.strides[kd]} + (col + {pl.cols[0] - B.origin[nd]}) '
               f'* {B.strides[nd]}')
        w(f'const {self.real} {v} = ({self.real})g_{B.arg.name}[{off}];')
        out[(t_idx, k)] = v
    return out

  # ---- compute ----
  def compute_block(self, bi, b0, acc):
    w = self.w
    pl = self.pl
    real = self.real
    b1 = self.blocks[bi][1]
    ring_slot = bi % 2
    block_seq = [(n, s) for n, s in enumerate(self.seq) if s[0] == bi]
    fused_terms = [ti for ti, t in enumerate(pl.terms) if t.kind == 'fused' and
                   (self.blocks[bi][2] is None or ti in self.blocks[bi][2])]
    offs = sorted(set(o for ti in fused_terms for _, o in self.slots[ti]))
    # sequence point: the block's LDS reads use offsets rewritten here, so the compiler cannot
    # hoist them above the previous block's FMAs (which would keep every operand live at once)
    w(f'asm volatile("" : {", ".join(f"{chr(34)}+v{chr(34)}({o})" for o in ["lane15"] + offs)});')
    chunks = {}
    touched = []
    for k in range(b0, b1):
      col_entries = [(n, s) for n, s in block_seq if s[2][1] == k]
      if not col_entries:
        continue
      rvals = {}
      terms_here = sorted(set(s[1] for _, s in col_entries))
      if any(pl.terms[ti].kind == 'fused' for ti in terms_here):
        qv = {}
        for o in offs:
          v = self.fresh('q')
          w(f'const {real} {v} = {self.ring_ref(ring_slot, o, k - b0)};')
          qv[o] = v
      for ti in terms_here:
        t = pl.terms[ti]
        if t.kind == 'fused':
          slots = self.slots[ti]
          v = self.fresh('r')
          w(f'{real} {v} = {qv[slots[0][1]]} * {slots[0][0]}_v;')
          for sv, o in slots[1:]:
            w(f'{v} = __builtin_fma({qv[o]}, {sv}_v, {v});')
          rvals[ti] = v
        else:
          rvals[ti] = self.global_rows[bi][(ti, k)]
      for n, (_, ti, (i, kk, pos)) in col_entries:
        c, j = divmod(n, CHUNK)
        if c not in chunks:
          cv = self.fresh('ch')
          w(f'const {real} {cv} = {self.chunk_ref(c)};')
          chunks[c] = cv
        self.fma(f'{acc}[{i}]', chunks[c], j, rvals[ti])
        if i not in touched:
          touched.append(i)
    for n0 in range(0, len(touched), 8):
      ops = ', '.join(f'"+v"({acc}[{i}])' for i in touched[n0:n0 + 8])
      w(f'asm volatile("" : {ops});')

  def compute_block_pipelined(self, bi, b0, acc):
    """compute_block with LDS reads issued ahead of their use: the shared-operand chunks
    ``stream_lds_lookahead`` chunks ahead, the streamed rows of step k+1 (and the fused
    temporary rows they give) during step k. Scheduling barriers keep the order, so the
    latency of each LDS read is covered by the FMAs issued in between."""
    w = self.w
    pl = self.pl
    real = self.real
    b1 = self.blocks[bi][1]
    ring_slot = bi % 2
    ahead = self.sched.stream_lds_lookahead
    block_seq = [(n, s) for n, s in enumerate(self.seq) if s[0] == bi]
    fused_terms = [ti for ti, t in enumerate(pl.terms) if t.kind == 'fused' and
                   (self.blocks[bi][2] is None or ti in self.blocks[bi][2])]
    offs = sorted(set(o for ti in fused_terms for _, o in self.slots[ti]))
    w(f'asm volatile("" : {", ".join(f"{chr(34)}+v{chr(34)}({o})" for o in ["lane15"] + offs)});')
    steps = []
    for k in range(b0, b1):
      ents = [(n, s) for n, s in block_seq if s[2][1] == k]
      if ents:
        steps.append((k, ents))
    if not steps:
      return
    order = []
    for _, ents in steps:
      for n, _ in ents:
        if n // CHUNK not in order:
          order.append(n // CHUNK)
    cvar = {}
    issued = [0]

    def issue(upto):
      while issued[0] < min(upto, len(order)):
        c = order[issued[0]]
        if c not in cvar:
          if self.pair:
            base = c - c % 2
            v = self.fresh('cp')
            w(f'const tf_d2 {v} = {self.chunk_pair_ref(base)};')
            cvar[base], cvar[base + 1] = f'{v}.x', f'{v}.y'
          else:
            v = self.fresh('ch')
            w(f'const {real} {v} = {self.chunk_ref(c)};')
            cvar[c] = v
        issued[0] += 1

    pend = {}
    pair_q = self.pair and pl.L is not None and self.ps % 2 == 0 and not self.dma

    def q_reads(k, ents):
      terms = set(s[1] for _, s in ents)
      if not any(pl.terms[ti].kind == 'fused' for ti in terms):
        return {}
      qv = {}
      kk = k - b0
      for o in offs:
        if pair_q and (o, kk) in pend:
          qv[o] = pend.pop((o, kk))
        elif pair_q and kk % 2 == 0 and k + 1 < b1:
          v = self.fresh('qp')
          w(f'const tf_d2 {v} = *reinterpret_cast<const tf_d2*>(&{self.ring_ref(ring_slot, o, kk)});')
          qv[o] = f'{v}.x'
          pend[(o, kk + 1)] = f'{v}.y'
        else:
          v = self.fresh('q')
          w(f'const {real} {v} = {self.ring_ref(ring_slot, o, kk)};')
          qv[o] = v
      return qv

    def t_rows(k, ents, qv):
      rv = {}
      for ti in sorted(set(s[1] for _, s in ents)):
        t = pl.terms[ti]
        if t.kind == 'fused':
          slots = self.slots[ti]
          v = self.fresh('r')
          w(f'{real} {v} = {qv[slots[0][1]]} * {slots[0][0]}_v;')
          for sv, o in slots[1:]:
            w(f'{v} = __builtin_fma({qv[o]}, {sv}_v, {v});')
          rv[ti] = v
        else:
          rv[ti] = self.global_rows[bi][(ti, k)]
      return rv

    barrier = '__builtin_amdgcn_sched_barrier(0);'
    issue(ahead)
    rvals = t_rows(steps[0][0], steps[0][1], q_reads(steps[0][0], steps[0][1]))
    touched = []
    for si, (k, ents) in enumerate(steps):
      nxt = steps[si + 1] if si + 1 < len(steps) else None
      w(barrier)
      nq = q_reads(nxt[0], nxt[1]) if nxt else {}
      groups = []
      for n, (_, ti, (i, kk, pos)) in ents:
        c, j = divmod(n, CHUNK)
        if not groups or groups[-1][0] != c:
          groups.append((c, []))
        groups[-1][1].append((ti, i, j))
      nrv = None
      for gi, (c, items) in enumerate(groups):
        issue(order.index(c) + 1 + ahead)
        w(barrier)
        if gi == min(1, len(groups) - 1) and nxt is not None:
          nrv = t_rows(nxt[0], nxt[1], nq)
        for ti, i, j in items:
          self.fma(f'{acc}[{i}]', cvar[c], j, rvals[ti])
          if i not in touched:
            touched.append(i)
      if nxt is not None:
        rvals = nrv
    w(barrier)
    for n0 in range(0, len(touched), 8):
      ops = ', '.join(f'"+v"({acc}[{i}])' for i in touched[n0:n0 + 8])
      w(f'asm volatile("" : {ops});')

  def fma(self, acc, chunk, lane_in_row, r):
    """acc += chunk[lane_in_row of the lane's row of 16] * r."""
    w = self.w
    ty = 'f64' if self.fp == FloatingPointType.DOUBLE else 'f32'
    if self.sched.dpp == 'asm':
      w(f'asm("v_fmac_{ty}_dpp %0, %1, %2 row_newbcast:{lane_in_row} row_mask:0xf bank_mask:0xf" '
        f': "+v"({acc}) : "v"({chunk}), "v"({r}));')
    else:
      b = self.fresh('bc')
      w(f'const {self.real} {b} = __builtin_amdgcn_mov_dpp({chunk}, {0x150 + lane_in_row}, 0xf, 0xf, true);')
      w(f'{acc} = __builtin_fma({b}, {r}, {acc});')


def _row_passes(pl, schedule):
  """Row ranges of the passes: one pass unless the destination has more rows than
  ``stream_pass_rows`` and its shared operands are dense (the runtime block loop)."""
  rows = pl.rows[1] - pl.rows[0]
  limit = schedule.stream_pass_rows
  if limit <= 0 or rows <= limit or any(t.U.is_sparse() for t in pl.terms):
    return [pl.rows]
  n = (rows + limit - 1) // limit
  step = (rows + n - 1) // n
  return [(r, min(r + step, pl.rows[1])) for r in range(pl.rows[0], pl.rows[1], step)]


def emit(program, context, base_name, schedule, pl) -> Emitted:
  passes = _row_passes(pl, schedule)
  if len(passes) == 1:
    return _Stream(program, context, base_name, schedule, pl).emit()
  if program.epilogue is not None:
    raise GenerationError('pointwise statements do not fuse into a chain run in row bands')
  # several passes over the contraction, each for a band of destination rows: the accumulator
  # of a band fits the registers and the band's shared rows fit LDS; the streamed operand is
  # read once per pass
  import dataclasses
  parts = []
  for r0, r1 in passes:
    sub = dataclasses.replace(pl, rows=(r0, r1),
                              terms=[dataclasses.replace(t, entries=[]) for t in pl.terms])
    parts.append(_Stream(program, context, base_name, schedule, sub).emit())
  first = parts[0]
  kname = first.kernel_name
  lines = first.kernel_src.rstrip('\n').split('\n')
  sig = next(i for i, l in enumerate(lines) if l.startswith('extern "C" __global__'))
  # every band's file-scope tables, then the kernel
  head = []
  for part in parts:
    plines = part.kernel_src.rstrip('\n').split('\n')
    psig = next(i for i, l in enumerate(plines) if l.startswith('extern "C" __global__'))
    head += plines[:psig]
  head += lines[sig:sig + 2]
  body = []
  for k, part in enumerate(parts):
    pl_lines = part.kernel_src.rstrip('\n').split('\n')
    psig = next(i for i, l in enumerate(pl_lines) if l.startswith('extern "C" __global__'))
    body.append(f'  // rows [{passes[k][0]}, {passes[k][1]})')
    body.append('  {')
    body += ['  ' + l for l in pl_lines[psig + 2:-1]]
    body.append('  }')
    if k + 1 < len(parts):
      body.append('  __syncthreads();')
  src = '\n'.join(head + body + ['}']) + '\n'
  lds = max(p.lds_bytes for p in parts)
  sched = dict(first.schedule)
  sched['row_passes'] = len(parts)
  sched['lds_bytes'] = lds
  return dataclasses.replace(first, kernel_src=src, lds_bytes=lds, schedule=sched,
                             extra_launch_setup=first.extra_launch_setup.replace(
                                 str(first.lds_bytes), str(lds)))
