"""The tuned gfx950 emitter for GEMM-shaped chains: one output column per lane.

Mapping. A workgroup is one wavefront of 64 lanes holding ``G = 64 // q`` elements, where ``q``
is the number of output columns of the chain's products (9 quantities for SeisSol's chains). Lane
``l`` owns column ``l % q`` of element ``l // q`` of every product's result and keeps it in
registers; lanes ``>= G * q`` only help with cooperative copies.

Per product ``out(:, j) = sum_k L(:, k) R(k, j)``:

* a left operand shared by all elements (a kernel argument with ``Addressing.NONE``, dense or
  sparse) or baked into the kernel is uniform across the wavefront: its entries are read with
  scalar loads (or are immediates) in their storage order, and each stored entry costs exactly one
  FMA per lane. Zeros of a sparse pattern cost nothing and no lane branches;
* a per-element left operand is staged once into LDS with coalesced copies, padded so the lanes'
  gathers spread over the banks, and read there;
* the right operand is read from the lane's own registers when it is the previous temporary,
  from global memory along the lane's column otherwise; a sparse per-element right operand is a
  per-lane list of (row, value) slots padded to the longest column of its pattern;
* results that later products accumulate into stay in registers until the last product, then are
  written once.

``Schedule`` holds the knobs the autotuner varies; ``unsupported_reason`` says why a chain falls
back to the plain emitter.
"""
import math
from dataclasses import asdict, dataclass, field
from typing import Dict, List, Optional, Tuple

from ..common.exceptions import GenerationError
from ..common.basic_types import Addressing, FloatingPointType
from . import ir
from .launcher import Emitted, Writer, kernel_params
from .traffic import element_traffic

WAVE = 64


@dataclass
class Schedule:
  """Tunable choices of the tuned emitter."""
  # traversal of a uniform left operand: 'storage' (its memory order, contiguous scalar loads)
  # or 'k_outer' (column by column, streams the right operand)
  uniform_order: str = 'storage'
  # rows of the output column processed per pass over the contraction (0 = all)
  row_block: int = 0
  # LDS leading-dim and per-element padding (in words of the compute type); None = auto
  lds_ld_pad: Optional[int] = None
  lds_elem_pad: Optional[int] = None
  # waves per workgroup (each wave owns its elements and LDS slice)
  waves_per_block: int = 1
  # read the right operand's column ahead of use ('column') or per step ('stream')
  right_load: str = 'auto'
  # contraction block of a uniform-left product (0 = whole contraction at once)
  k_block: int = 16
  # uniform-operand software pipeline: entries per chunk (0 = off) and chunks loaded ahead
  uniform_chunk: int = 0
  uniform_lookahead: int = 1
  # scheduling fence after each contraction block
  block_fence: bool = True
  # give each block's scalar loads an opaque base so loads of different blocks are not merged
  opaque_block_base: bool = True
  # pin each block's results in registers (stops the compiler sinking the block's FMAs)
  pin_results: bool = True
  # compute a temporary consumed once, as a right operand, row block by row block inside its
  # consumer instead of materialising it
  fuse_temporaries: bool = True
  # minimum waves per SIMD requested from the register allocator (launch bounds hint)
  min_waves_per_simd: int = 1
  # workgroup barrier before each uniform-left product: keeps the wavefronts of a workgroup in
  # the same phase so they share the scalar cache's copy of the uniform operand
  phase_sync: bool = False
  # streaming schedule (persistent workgroups, shared operands in LDS, DPP-broadcast FMAs) when
  # the chain has its form; its own workgroup size and contraction block
  stream: bool = True
  stream_waves: int = 8
  stream_k_block: int = 8
  # contraction block of the runtime block loop (dense shared operands with many entries)
  stream_loop_k_block: int = 4
  # destination rows per pass of the runtime block loop (0: one pass)
  stream_pass_rows: int = 64
  stream_min_waves: int = 2
  # runtime loop over contraction blocks: 'auto' (dense shared operands with many entries),
  # 'always', 'never'; where the shared entries are read from: 'lds', 'global', 'auto'
  stream_loop: str = 'auto'
  stream_u_source: str = 'auto'
  # contraction blocks of the streamed operand loaded ahead of use (across element groups)
  stream_prefetch: int = 1
  # lane (element, column) streams its own column of the per-element operand
  stream_own_columns: bool = True
  # lanes load consecutive rows of one column from a uniform base (coalesced row blocks)
  stream_row_loads: bool = False
  # two shared-operand chunks (and two streamed rows) per LDS read
  stream_pair_reads: bool = False
  # shared-operand chunks loaded ahead of use (0: loads placed by the compiler)
  stream_lds_lookahead: int = 2
  # rows of the streamed operand loaded straight into LDS (global_load_lds), ring of this many
  # blocks (prefetch distance = slots - 1)
  stream_dma: bool = False
  stream_dma_slots: int = 3
  # how the DPP broadcast is issued: 'asm' (one v_fmac_*_dpp) or 'builtin' (v_mov_*_dpp + fma)
  dpp: str = 'asm'
  # FP64 matrix-core form: wavefronts per workgroup and contraction rows per block (0: chosen
  # from the chain), and where the destination's old values wait for the accumulation:
  # 'acc' (loaded into the accumulators before the first MFMA), 'regs' (loaded into registers of
  # their own at the start of the group, added at the end), 'end' (loaded at the end), '' (chosen)
  mfma_waves: int = 0
  mfma_k_block: int = 0
  mfma_dest: str = ''

  def key(self):
    return tuple(sorted(asdict(self).items()))


def _real(fp):
  return FloatingPointType.as_str(fp)


def _lit(fp, v):
  return fp.literal(v)


@dataclass
class _ProductShape:
  left: ir.Factor
  right: ir.Factor
  left_dims: Tuple[int, int]      # storage dim of (n0, k) in the left factor
  right_dims: Tuple[int, int]     # storage dim of (k, n1) in the right factor


def _classify(p: ir.Product):
  """Return (_ProductShape, None) or (None, reason)."""
  if p.out_rank != 2:
    return None, f'product of rank {p.out_rank} (only rank-2 products are tuned)'
  if len(p.k_ranges) != 1:
    return None, f'{len(p.k_ranges)} contracted indices (only one is tuned)'
  if len(p.factors) != 2:
    return None, f'{len(p.factors)} matrix operands (only two are tuned)'
  left = right = None
  for f in p.factors:
    labs = set(f.labels)
    if labs == {('n', 0), ('k', 0)} and f.rank == 2:
      left = f
    elif labs == {('k', 0), ('n', 1)} and f.rank == 2:
      right = f
  if left is None or right is None:
    return None, 'operands are not of the form L(n0,k) R(k,n1)'
  ld = (left.labels.index(('n', 0)), left.labels.index(('k', 0)))
  rd = (right.labels.index(('k', 0)), right.labels.index(('n', 1)))
  return _ProductShape(left, right, ld, rd), None


def _left_kind(f: ir.Factor):
  if f.src == ir.SRC_DATA:
    return 'data'
  if f.src == ir.SRC_GLOBAL and f.arg.addressing == Addressing.NONE:
    return 'uniform'
  if f.src == ir.SRC_GLOBAL and f.arg.per_element:
    return 'lds'
  if f.src == ir.SRC_TMP:
    return 'lds_tmp'
  return None


def unsupported_reason(program: ir.ChainProgram, context, schedule=None):
  ncols = None
  tmp_cols = {}
  for p in program.products:
    if p.is_empty():
      return 'empty product'
    shape, why = _classify(p)
    if shape is None:
      return why
    lo1, hi1 = p.n_ranges[1]
    q = hi1 - lo1
    if ncols is None:
      ncols = q
    elif q != ncols:
      return 'products with different column counts'
    if q > WAVE // 2:
      return f'{q} output columns (at most {WAVE // 2} are tuned)'
    if p.n_ranges[0][1] - p.n_ranges[0][0] > 96:
      return 'more than 96 output rows'
    lk = _left_kind(shape.left)
    if lk is None:
      return f'left operand kind {shape.left.src}'
    if lk == 'lds' and shape.left.is_sparse():
      return 'sparse per-element left operand'
    if lk in ('lds', 'lds_tmp') and shape.right.src == ir.SRC_GLOBAL and \
        shape.right.is_sparse() and shape.left_dims != (0, 1):
      return 'transposed per-element left operand with a sparse right operand'
    r = shape.right
    if r.src == ir.SRC_TMP:
      if shape.right_dims != (0, 1) or lo1 != 0:
        return 'transposed or offset temporary as right operand'
    elif r.src == ir.SRC_GLOBAL:
      pass
    else:
      return f'right operand kind {r.src}'
    if shape.left.src == ir.SRC_TMP and shape.left_dims != (0, 1):
      return 'transposed temporary as left operand'
    for s in p.scalars:
      if s.src not in (ir.SRC_DATA, ir.SRC_SCALAR):
        return 'unsupported scalar'
    if p.dest_src == ir.SRC_TMP and lo1 != 0:
      return 'temporary with offset columns'
  if ncols is None:
    return 'no products'
  # a global destination read back as an operand later in the chain
  written = set()
  for p in program.products:
    for f in p.factors:
      if f.src == ir.SRC_GLOBAL and f.arg.name in written:
        return f'{f.arg.name} is read after the chain writes it'
    if p.dest_src == ir.SRC_GLOBAL:
      written.add(p.dest_arg.name)
  return None


# ------------------------------------------------------------------------- emission

class _Emitter:
  def __init__(self, program, context, base_name, schedule):
    self.prog = program
    self.ctx = context
    self.fp = program.fp_type
    self.real = _real(self.fp)
    self.words = self.fp.size() // 4       # 4-byte LDS words per value
    self.sched = schedule or Schedule()
    self.kname = f'kernel_{base_name}'
    self.w = Writer()
    self.q = program.products[0].n_ranges[1][1] - program.products[0].n_ranges[1][0]
    self.G = WAVE // self.q
    self.counter = 0
    self.lds_regions: Dict[Tuple, dict] = {}
    self.lds_words_per_elem = 0
    self.tmp_regs: Dict[int, str] = {}
    self.dest_regs: Dict[str, dict] = {}
    self.fused: Dict[int, dict] = {}
    self.fused_producers = self._find_fusable() if self.sched.fuse_temporaries else {}

  def fresh(self, prefix):
    self.counter += 1
    return f'{prefix}{self.counter}'

  # ---- LDS layout ----
  def plan_lds(self):
    """Regions for per-element left operands (global windows and temporaries), with paddings
    chosen so the lanes' gathers hit distinct banks."""
    offset = 0
    for pi, p in enumerate(self.prog.products):
      shape, _ = _classify(p)
      kind = _left_kind(shape.left)
      if kind not in ('lds', 'lds_tmp'):
        continue
      key = self._lds_key(p, shape)
      if key in self.lds_regions:
        continue
      rows = p.n_ranges[0][1] - p.n_ranges[0][0]
      kr = p.k_ranges[0]
      if kind == 'lds':
        f = shape.left
        win_rows = (p.n_ranges[0][0], p.n_ranges[0][1])
        win_cols = kr
        # store the window with its leading dim along the left's storage dim 0
        d0, d1 = 0, 1
        r0 = win_rows if shape.left_dims[0] == 0 else win_cols
        r1 = win_cols if shape.left_dims[0] == 0 else win_rows
        ext0, ext1 = r0[1] - r0[0], r1[1] - r1[0]
      else:
        tshape = self.prog.tmp_shapes[id(shape.left.tensor)]
        r0, r1 = (0, tshape[0]), (0, tshape[1])
        ext0, ext1 = tshape
      ld = self._pick_ld(ext0, ext1, p, shape)
      self.lds_regions[key] = {'offset': offset, 'ld': ld, 'ext0': ext0, 'ext1': ext1, 'r0': r0,
                               'r1': r1, 'kind': kind, 'factor': shape.left, 'filled': False}
      offset += ld * ext1
    self.lds_words_per_elem_values = offset
    elem = offset
    pad = self.sched.lds_elem_pad
    if pad is None:
      pad = self._pick_elem_pad(elem)
    self.elem_stride = elem + pad if elem else 0

  def _lds_key(self, p, shape):
    f = shape.left
    if f.src == ir.SRC_TMP:
      return ('tmp', id(f.tensor))
    return ('glb', f.arg.name, tuple(p.n_ranges[0]), tuple(p.k_ranges[0]), shape.left_dims)

  def _pick_ld(self, ext0, ext1, p, shape):
    pad = self.sched.lds_ld_pad
    if pad is not None:
      return ext0 + pad
    return ext0 + (1 if ext0 % 2 == 0 else 0)

  def _pick_elem_pad(self, elem):
    # spread element slots over banks: (elem + pad) values, odd in values
    return 0 if elem % 2 == 1 else 1

  # ---- kernel ----
  def emit(self) -> Emitted:
    self.plan_lds()
    w = self.w
    fp, real = self.fp, self.real
    waves = max(1, self.sched.waves_per_block)
    threads = WAVE * waves
    params = ', '.join(kernel_params(self.prog, self.ctx))
    mw = max(1, self.sched.min_waves_per_simd)
    w(f'extern "C" __global__ void __launch_bounds__({threads}, {mw}) {self.kname}({params})')
    w('{')
    w._depth += 1
    w(f'// tuned emitter: {self.G} elements x {self.q} columns per wavefront, '
      f'{waves} wavefront(s) per block; schedule {asdict(self.sched)}')
    lds_total = self.elem_stride * self.G * waves
    if lds_total:
      w(f'extern __shared__ __align__(16) {real} lds_all[];')
      w(f'{real}* __restrict__ lds = lds_all + {self.elem_stride * self.G} * (threadIdx.x / {WAVE});')
    w(f'const int lane = threadIdx.x % {WAVE};')
    # lanes past the last element slot work on a copy of the last slot's data (results unused)
    w(f'const int slot = lane / {self.q} < {self.G} ? lane / {self.q} : {self.G - 1};')
    w(f'const int col = lane / {self.q} < {self.G} ? lane - slot * {self.q} : 0;')
    w(f'const size_t e0 = ((size_t)blockIdx.x * {waves} + threadIdx.x / {WAVE}) * {self.G};')
    w('const size_t batchId = e0 + slot;')
    w(f'const bool active = (lane < {self.G * self.q}) && (batchId < numElements) && '
      '(flags == nullptr || flags[batchId] != 0u);')
    # Lanes without an element of their own address element 0 (it exists: the launcher does not
    # launch for numElements == 0), so every load below is unconditional and the hot code has
    # no branches; only the final stores test `active`.
    w('const size_t eSafe = active ? batchId : 0;')
    for a in self.prog.args:
      if a.kind != ir.SRC_GLOBAL:
        continue
      ty = f'{"const " if a.const else ""}{a.dtype}'
      if a.addressing == Addressing.NONE:
        w(f'{ty}* __restrict__ g_{a.name} = {a.name};')
      elif a.addressing == Addressing.STRIDED:
        w(f'{ty}* __restrict__ g_{a.name} = {a.name} + eSafe * {a.element_stride}ull + '
          f'{a.name}_extraOffset;')
      else:
        w(f'{ty}* __restrict__ g_{a.name} = {a.name}[eSafe] + {a.name}_extraOffset;')
    for pi, p in enumerate(self.prog.products):
      self.emit_product(pi, p)
    self.flush_dests()
    w._depth -= 1
    w('}')
    traffic = element_traffic(self.prog)
    lds_bytes = lds_total * self.fp.size()
    return Emitted(kernel_name=self.kname, kernel_src=w.src(), block=(threads, 1, 1),
                   elements_per_block=self.G * waves, lds_bytes=lds_bytes, emitter='tuned',
                   schedule=dict(asdict(self.sched), elements_per_wave=self.G, columns=self.q,
                                 lds_words_per_element=self.elem_stride),
                   bytes_per_element=traffic.per_element, shared_bytes=traffic.shared)

  # ---- fusion of temporaries ----
  def _find_fusable(self):
    """Products whose temporary result is consumed once, as the right operand of a later
    uniform-left product, and whose rows can be computed independently from LDS: index ->
    consumer index."""
    out = {}
    prods = self.prog.products
    for a, p1 in enumerate(prods):
      if p1.dest_src != ir.SRC_TMP or p1.add:
        continue
      sh1, _ = _classify(p1)
      if _left_kind(sh1.left) != 'lds' or sh1.left_dims != (0, 1):
        continue
      if sh1.right.src != ir.SRC_GLOBAL or not sh1.right.arg.per_element:
        continue
      uses = [(b, f) for b, p2 in enumerate(prods) for f in p2.factors if f.tensor is p1.dest]
      writes = [b for b, p2 in enumerate(prods) if p2.dest is p1.dest and b != a]
      if len(uses) != 1 or writes:
        continue
      b, f = uses[0]
      if b <= a:
        continue
      sh2, _ = _classify(prods[b])
      if sh2.right is not f or sh2.right_dims != (0, 1) or prods[b].n_ranges[1][0] != 0:
        continue
      if _left_kind(sh2.left) not in ('uniform', 'data'):
        continue
      out[a] = b
    return out

  def setup_fused(self, pi, p, shape):
    """Stage the producer's left operand and load the lane's right-operand values; rows of the
    temporary are then formed on demand by :meth:`fused_row`."""
    w = self.w
    reg = self.stage_left(p, shape)
    es = self.elem_stride
    base = f'lds + slot * {es} + {reg["offset"]}'
    klo, khi = p.k_ranges[0]
    lo1 = p.n_ranges[1][0]
    R = shape.right
    rk, rn = shape.right_dims
    terms = []   # (lds column pointer var, value var)
    w(f'// temporary {p.dest.name} is formed row by row where it is consumed')
    if R.is_sparse():
      cols = []
      for c in range(self.q):
        ks = []
        for k in range(klo, khi):
          idx = [0, 0]
          idx[rk], idx[rn] = k, lo1 + c
          pos = R.packed.get(tuple(idx))
          if pos is not None:
            ks.append((k, pos))
        cols.append(ks)
      slots = max(len(c) for c in cols)
      for s_ in range(slots):
        ks = [c[s_][0] if s_ < len(c) else (c[0][0] if c else klo) for c in cols]
        ps = [c[s_][1] if s_ < len(c) else -1 for c in cols]
        kv = self.fresh('sk')
        w(f'const int {kv} = {self._select_int(ks)};')
        w(f'const int {kv}_p = {self._select_int(ps)};')
        w(f'const {self.real} {kv}_x = ({self.real})g_{R.arg.name}[{kv}_p < 0 ? 0 : {kv}_p];')
        w(f'const {self.real} {kv}_v = {kv}_p >= 0 ? {kv}_x : {_lit(self.fp, 0)};')
        w(f'unsigned {kv}_o = slot * {es} + {reg["offset"]} + ({kv} - {reg["r1"][0]}) * {reg["ld"]};')
        terms.append((f'{kv}_o', f'{kv}_v'))
    else:
      for k in range(klo, khi):
        kv = self.fresh('dk')
        off = (f'({k - R.origin[rk]}) * {R.strides[rk]} + (col + {lo1 - R.origin[rn]}) * '
               f'{R.strides[rn]}')
        w(f'const {self.real} {kv}_v = ({self.real})g_{R.arg.name}[{off}];')
        w(f'unsigned {kv}_o = slot * {es} + {reg["offset"] + (k - reg["r1"][0]) * reg["ld"]};')
        terms.append((f'{kv}_o', f'{kv}_v'))
    scale = self._scalar_expr(p)
    self.fused[id(p.dest)] = {'terms': terms, 'row0': p.n_ranges[0][0] - reg['r0'][0],
                              'scale': scale, 'rows': p.n_ranges[0][1] - p.n_ranges[0][0]}

  def fused_row(self, tid, k, cache):
    key = ('f', tid, k)
    if key in cache:
      return cache[key]
    info = self.fused[tid]
    if info['scale'] == 'ZERO' or not info['terms'] or not 0 <= k < info['rows']:
      cache[key] = None
      return None
    v = self.fresh('t')
    w = self.w
    a0, v0 = info['terms'][0]
    w(f'{self.real} {v} = lds[{a0} + {info["row0"] + k}] * {v0};')
    for a, vv in info['terms'][1:]:
      w(f'{v} = __builtin_fma(lds[{a} + {info["row0"] + k}], {vv}, {v});')
    if info['scale'] is not None:
      w(f'{v} = {info["scale"]} * {v};')
    cache[key] = v
    return v

  # ---- staging ----
  def stage_left(self, p, shape):
    key = self._lds_key(p, shape)
    reg = self.lds_regions[key]
    if reg['filled']:
      return reg
    w = self.w
    f = shape.left
    es = self.elem_stride
    if reg['kind'] == 'lds_tmp':
      treg = self.tmp_regs[id(f.tensor)]
      ext0 = reg['ext0']
      w(f'// stage temporary {f.tensor.name} into LDS')
      with w.block(f'if (lane < {self.G * self.q})'):
        for i in range(ext0):
          w(f'lds[slot * {es} + {reg["offset"]} + col * {reg["ld"]} + {i}] = {treg}[{i}];')
      w('__syncthreads();')
      reg['filled'] = True
      return reg
    a = f.arg
    ext0, ext1, ld = reg['ext0'], reg['ext1'], reg['ld']
    r0, r1 = reg['r0'], reg['r1']
    count = ext0 * ext1
    w(f'// stage {a.name} window {r0}x{r1} of {self.G} elements into LDS')
    vol = a.element_stride
    full = (r0 == (f.origin[0], f.origin[0] + f.tensor.shape[0]) and
            r1 == (f.origin[1], f.origin[1] + f.tensor.shape[1]) and not f.is_sparse())
    if f.is_sparse():
      raise NotImplementedError('sparse per-element left operand')
    if a.addressing == Addressing.STRIDED and full:
      # the G elements are contiguous: copy G*vol values with all lanes
      total = self.G * vol
      with w.block(''):
        w(f'const {a.dtype}* __restrict__ src = {a.name} + e0 * {vol}ull + {a.name}_extraOffset;')
        w(f'const int nvalid = (int)((numElements > e0) ? ((numElements - e0) < {self.G} ? '
          f'(numElements - e0) : {self.G}) : 0) * {vol};')
        with w.block(f'for (int i = lane; i < nvalid; i += {WAVE})'):
          w(f'const int s = i / {vol};')
          w(f'const int r = i - s * {vol};')
          w(f'const int c = r / {f.tensor.shape[0]};')
          w(f'const int rr = r - c * {f.tensor.shape[0]};')
          w(f'lds[s * {es} + {reg["offset"]} + c * {ld} + rr] = src[i];')
    else:
      with w.block(''):
        with w.block(f'for (int i = lane; i < {self.G * count}; i += {WAVE})'):
          w(f'const int s = i / {count};')
          w(f'const int r = i - s * {count};')
          w(f'const int c = r / {ext0};')
          w(f'const int rr = r - c * {ext0};')
          w(f'const size_t e = e0 + s;')
          with w.block('if (e < numElements)'):
            goff = (f'(rr + {r0[0] - f.origin[0]}) * {f.strides[0]} + '
                    f'(c + {r1[0] - f.origin[1]}) * {f.strides[1]}')
            if a.addressing == Addressing.STRIDED:
              src = f'{a.name}[e * {vol}ull + {a.name}_extraOffset + {goff}]'
            else:
              src = f'{a.name}[e][{a.name}_extraOffset + {goff}]'
            w(f'lds[s * {es} + {reg["offset"]} + c * {ld} + rr] = {src};')
    w('__syncthreads();')
    reg['filled'] = True
    return reg

  # ---- right operand access (per lane, its column) ----
  def right_value(self, p, shape, k, cache):
    """C expression of R(k, jcol) for this lane; None for a structural zero of every lane."""
    r = shape.right
    if r.src == ir.SRC_TMP and id(r.tensor) in self.fused:
      return self.fused_row(id(r.tensor), k, cache)
    if r.src == ir.SRC_TMP:
      return f'{self.tmp_regs[id(r.tensor)]}[{k}]'
    key = ('r', k)
    if key in cache:
      return cache[key]
    a = r.arg
    lo1 = p.n_ranges[1][0]
    kd, nd = shape.right_dims
    var = self.fresh('rv')
    w = self.w
    if r.is_sparse():
      # per-lane lookup of (k, jcol) in the pattern: a compile-time table per column
      positions = []
      any_nz = False
      for c in range(self.q):
        idx = [0, 0]
        idx[kd], idx[nd] = k, lo1 + c
        pos = r.packed.get(tuple(idx))
        positions.append(-1 if pos is None else pos)
        any_nz |= pos is not None
      if not any_nz:
        cache[key] = None
        return None
      sel = self._select_int(positions)
      w(f'const int {var}_p = {sel};')
      w(f'const {self.real} {var}_x = ({self.real})g_{a.name}[{var}_p < 0 ? 0 : {var}_p];')
      w(f'const {self.real} {var} = {var}_p >= 0 ? {var}_x : {_lit(self.fp, 0)};')
    else:
      off = f'({k - r.origin[kd]}) * {r.strides[kd]} + (col + {lo1 - r.origin[nd]}) * {r.strides[nd]}'
      w(f'const {self.real} {var} = ({self.real})g_{a.name}[{off}];')
    cache[key] = var
    return var

  def _pin(self, values, per_stmt=8):
    """Empty asm that reads and rewrites ``values``: the FMAs producing them must run before
    it, so the compiler cannot sink a block's arithmetic past later blocks (which would keep
    every loaded operand alive at once)."""
    for n in range(0, len(values), per_stmt):
      group = values[n:n + per_stmt]
      ops = ', '.join(f'"+v"({v})' for v in group)
      self.w(f'asm volatile("" : {ops});')

  def _select_int(self, values):
    """C expression choosing values[col] (compile-time table, lane-dependent index)."""
    if all(v == values[0] for v in values):
      return str(values[0])
    expr = str(values[-1])
    for c in range(len(values) - 2, -1, -1):
      expr = f'(col == {c} ? {values[c]} : {expr})'
    return expr

  # ---- products ----
  def dest_acc(self, p, pi):
    """Register array the product writes, and its state: 'new' (no value yet), 'loaded'
    (old destination times beta) or 'existing' (an earlier product's value)."""
    rows = p.n_ranges[0][1] - p.n_ranges[0][0]
    w = self.w
    if p.dest_src == ir.SRC_TMP:
      tid = id(p.dest)
      if tid in self.tmp_regs:
        return self.tmp_regs[tid], 'existing', rows
      name = f'tmp_{p.dest.name}'
      self.tmp_regs[tid] = name
      w(f'{self.real} {name}[{rows}];')
      return name, 'new', rows
    key = p.dest_arg.name
    if key in self.dest_regs:
      d = self.dest_regs[key]
      if d['range'] != tuple(p.n_ranges[0]):
        raise NotImplementedError('destination written with different row ranges')
      return d['name'], 'existing', rows
    name = f'acc_{key}'
    self.dest_regs[key] = {'name': name, 'range': tuple(p.n_ranges[0]), 'product': p}
    w(f'{self.real} {name}[{rows}];')
    if p.add:
      lo0 = p.n_ranges[0][0]
      lo1 = p.n_ranges[1][0]
      for i in range(rows):
        off = self._dest_off(p, lo0 + i, lo1)
        old = f'g_{key}[{off}]'
        val = old if p.beta == 1.0 else f'{_lit(self.fp, p.beta)} * {old}'
        w(f'{name}[{i}] = ({self.real})({val});')
      return name, 'loaded', rows
    return name, 'new', rows

  def _dest_off(self, p, row, lo1):
    o0, o1 = p.dest_origin
    s0, s1 = p.dest_strides
    return f'{(row - o0) * s0} + (col + {lo1 - o1}) * {s1}'

  def emit_product(self, pi, p):
    w = self.w
    shape, _ = _classify(p)
    w()
    w(f'// product {pi}: {p.descr}')
    kind = _left_kind(shape.left)
    if pi in self.fused_producers:
      self.setup_fused(pi, p, shape)
      return
    if kind in ('lds', 'lds_tmp'):
      self.stage_left(p, shape)
    if self.sched.phase_sync and kind in ('uniform', 'data'):
      w('__syncthreads();')
    scale = self._scalar_expr(p)
    acc, state, rows = self.dest_acc(p, pi)
    zero = _lit(self.fp, 0)
    if not p.add:
      for i in range(rows):
        w(f'{acc}[{i}] = {zero};')
      if scale != 'ZERO':
        self._accumulate(p, shape, kind, acc)
        if scale is not None:
          for i in range(rows):
            w(f'{acc}[{i}] = {scale} * {acc}[{i}];')
      return
    if state == 'existing' and p.beta != 1.0:
      for i in range(rows):
        w(f'{acc}[{i}] = {_lit(self.fp, p.beta)} * {acc}[{i}];')
    if scale == 'ZERO':
      return
    if scale is None:
      self._accumulate(p, shape, kind, acc)
      return
    part = self.fresh('part')
    w(f'{self.real} {part}[{rows}];')
    for i in range(rows):
      w(f'{part}[{i}] = {zero};')
    self._accumulate(p, shape, kind, part)
    for i in range(rows):
      w(f'{acc}[{i}] = {acc}[{i}] + {scale} * {part}[{i}];')

  def _accumulate(self, p, shape, kind, target):
    if kind in ('uniform', 'data'):
      self.uniform_left(p, shape, target)
    else:
      self.lds_left(p, shape, target)

  def _scalar_expr(self, p):
    parts = []
    for s in p.scalars:
      if s.src == ir.SRC_DATA:
        v = next(iter(s.values.values()), None) if s.values else None
        if not v:
          return 'ZERO'
        if v != 1.0:
          parts.append(_lit(self.fp, v))
      else:
        parts.append(f'(({self.real}){s.arg.name})')
    return ' * '.join(parts) if parts else None

  def uniform_left(self, p, shape, target):
    """out(:, j) += L(:, k) R(k, j) with L uniform: one FMA per stored entry of L.

    The contraction is cut into blocks of ``k_block`` consecutive k. Per block the right
    operand's rows are made available (registers, streamed loads or rows of a fused temporary),
    then the block's stored entries of L are visited in storage order, so each row's entries of a
    CSR operand (or each column's of a column-major one) are contiguous scalar loads."""
    w = self.w
    L = shape.left
    lo0 = p.n_ranges[0][0]
    klo, khi = p.k_ranges[0]
    rows = p.n_ranges[0][1] - lo0
    nd, kd = shape.left_dims
    entries = []   # (position, row i relative, k, value expression)
    if L.src == ir.SRC_DATA:
      for n, (idx, v) in enumerate(sorted(L.values.items(), key=lambda kv: (kv[0][kd], kv[0][nd]))):
        i, k = idx[nd], idx[kd]
        if v != 0.0 and lo0 <= i < lo0 + rows and klo <= k < khi:
          entries.append((n, i - lo0, k, _lit(self.fp, v)))
    else:
      if L.is_sparse():
        items = sorted(L.packed.items(), key=lambda kv: kv[1])
      else:
        items = [(idx, L.offset(idx)) for idx in
                 ir.iterate_box([(L.origin[d], L.origin[d] + L.tensor.shape[d]) for d in range(2)])]
        items.sort(key=lambda kv: kv[1])
      for idx, pos in items:
        i, k = idx[nd], idx[kd]
        if lo0 <= i < lo0 + rows and klo <= k < khi:
          entries.append((pos, i - lo0, k, None))
    kb = self.sched.k_block if self.sched.k_block > 0 else (khi - klo)
    chunk = self.sched.uniform_chunk if L.src != ir.SRC_DATA else 0
    ahead = max(1, self.sched.uniform_lookahead)
    for b0 in range(klo, khi, kb):
      block = [e for e in entries if b0 <= e[2] < b0 + kb]
      if not block:
        continue
      if self.sched.uniform_order == 'k_outer':
        block.sort(key=lambda e: (e[2], e[1]))
      else:
        block.sort(key=lambda e: e[0])
      w(f'// k block [{b0}, {min(b0 + kb, khi)})')
      if shape.right.src == ir.SRC_TMP and id(shape.right.tensor) in self.fused and \
          self.sched.opaque_block_base:
        # the block's LDS reads of the fused temporary's rows depend on offsets "rewritten"
        # here, so they cannot be hoisted into earlier blocks
        for off, _ in self.fused[id(shape.right.tensor)]['terms']:
          w(f'asm volatile("" : "+v"({off}));')
      if L.src != ir.SRC_DATA:
        # Scalar loads of this block go through a base pointer the compiler cannot relate to
        # the other blocks' (an opaque zero offset), so it cannot merge a row's entries of
        # different blocks into one wide scalar load hoisted far ahead of their use.
        bvar = self.fresh('kb')
        if self.sched.opaque_block_base:
          w(f'unsigned {bvar}_z; asm volatile("s_mov_b32 %0, 0" : "=s"({bvar}_z));')
          w(f'const {L.arg.dtype}* __restrict__ {bvar} = g_{L.arg.name} + {bvar}_z;')
        else:
          w(f'const {L.arg.dtype}* __restrict__ {bvar} = g_{L.arg.name};')
        block = [(pos, i, k, f'{bvar}[{pos}]') for pos, i, k, _ in block]
      cache = {}
      for k in sorted({e[2] for e in block}):
        self.right_value(p, shape, k, cache)
      if chunk <= 0:
        touched = []
        for _, i, k, val in block:
          r = self.right_value(p, shape, k, cache)
          if r is not None:
            w(f'{target}[{i}] = __builtin_fma(({self.real}){val}, {r}, {target}[{i}]);')
            if i not in touched:
              touched.append(i)
        if self.sched.pin_results:
          self._pin([f'{target}[{i}]' for i in touched])
        if self.sched.block_fence:
          w('__builtin_amdgcn_sched_barrier(0);')
        continue
      # software pipeline: the uniform values of chunk c + ahead load while chunk c computes;
      # fences stop the compiler from hoisting every scalar load to the top
      chunks = [block[n:n + chunk] for n in range(0, len(block), chunk)]
      names = {}

      def load_chunk(c):
        for n, (_, i, k, val) in enumerate(chunks[c]):
          v = self.fresh('u')
          names[(c, n)] = v
          w(f'const {self.real} {v} = ({self.real}){val};')

      for c in range(min(ahead, len(chunks))):
        load_chunk(c)
      for c in range(len(chunks)):
        w('__builtin_amdgcn_sched_barrier(0);')
        if c + ahead < len(chunks):
          load_chunk(c + ahead)
        for n, (_, i, k, val) in enumerate(chunks[c]):
          r = self.right_value(p, shape, k, cache)
          if r is not None:
            w(f'{target}[{i}] = __builtin_fma({names[(c, n)]}, {r}, {target}[{i}]);')
      w('__builtin_amdgcn_sched_barrier(0);')

  def lds_left(self, p, shape, target):
    """out(:, j) += L(:, k) R(k, j) with L per element, staged in LDS."""
    w = self.w
    reg = self.lds_regions[self._lds_key(p, shape)]
    es = self.elem_stride
    lo0 = p.n_ranges[0][0]
    rows = p.n_ranges[0][1] - lo0
    klo, khi = p.k_ranges[0]
    nd, kd = shape.left_dims
    L = shape.left
    R = shape.right
    base = f'lds + slot * {es} + {reg["offset"]}'

    if R.src == ir.SRC_GLOBAL and R.is_sparse():
      # per-lane slot lists over the stored rows k of the lane's column, padded to the longest
      lo1 = p.n_ranges[1][0]
      rk, rn = shape.right_dims
      cols = []
      for c in range(self.q):
        ks = []
        for k in range(klo, khi):
          idx = [0, 0]
          idx[rk], idx[rn] = k, lo1 + c
          pos = R.packed.get(tuple(idx))
          if pos is not None:
            ks.append((k, pos))
        cols.append(ks)
      slots = max(len(c) for c in cols)
      slot_vars = []
      for s in range(slots):
        ks = [c[s][0] if s < len(c) else (c[0][0] if c else klo) for c in cols]
        ps = [c[s][1] if s < len(c) else -1 for c in cols]
        kv = self.fresh('sk')
        w(f'const int {kv} = {self._select_int(ks)};')
        w(f'const int {kv}_p = {self._select_int(ps)};')
        w(f'const {self.real} {kv}_x = ({self.real})g_{R.arg.name}[{kv}_p < 0 ? 0 : {kv}_p];')
        w(f'const {self.real} {kv}_v = {kv}_p >= 0 ? {kv}_x : {_lit(self.fp, 0)};')
        w(f'const {self.real}* __restrict__ {kv}_a = {self._lds_col_ptr(reg, shape, kv, base)};')
        slot_vars.append(kv)
      for i in range(rows):
        ia = lo0 + i
        for sv in slot_vars:
          w(f'{target}[{i}] = __builtin_fma({sv}_a[{ia - reg["r0"][0]}], {sv}_v, {target}[{i}]);')
      return
    # dense right operand: one LDS read of L(i,k) per term
    cache = {}
    for k in range(klo, khi):
      r = self.right_value(p, shape, k, cache)
      if r is None:
        continue
      for i in range(rows):
        ia = lo0 + i
        w(f'{target}[{i}] = __builtin_fma(({base})[{self._elem_off(reg, shape, ia, k)}], {r}, '
          f'{target}[{i}]);')

  def _lds_col_ptr(self, reg, shape, kvar, base):
    """Pointer to L(r0_lo, k) for a per-lane k (left stored with rows along dim 0)."""
    nd, kd = shape.left_dims
    if nd != 0:
      raise NotImplementedError('transposed per-element left with sparse right')
    return f'({base}) + ({kvar} - {reg["r1"][0]}) * {reg["ld"]}'

  def _elem_off(self, reg, shape, ia, k):
    nd, kd = shape.left_dims
    if nd == 0:
      return (ia - reg['r0'][0]) + (k - reg['r1'][0]) * reg['ld']
    return (k - reg['r0'][0]) + (ia - reg['r1'][0]) * reg['ld']

  def flush_dests(self):
    w = self.w
    for key, d in self.dest_regs.items():
      p = d['product']
      lo0, hi0 = d['range']
      lo1 = p.n_ranges[1][0]
      w()
      w(f'// store {key}')
      with w.block('if (active)'):
        for i in range(hi0 - lo0):
          w(f'g_{key}[{self._dest_off(p, lo0 + i, lo1)}] = {d["name"]}[{i}];')


def starved(emitted) -> bool:
  """A column-per-lane kernel whose LDS per workgroup leaves fewer than four wavefronts
  resident on a CU (one per SIMD): nothing hides its latencies, and the plain kernel, with
  less LDS per element, runs faster."""
  if emitted.schedule.get('kind') == 'stream' or not emitted.lds_bytes:
    return False
  waves = emitted.block[0] * emitted.block[1] * emitted.block[2] // 64
  return (LDS_LIMIT // emitted.lds_bytes) * waves < 4


def emit(program, context, base_name, schedule=None) -> Emitted:
  from . import stream
  schedule = schedule or Schedule()
  if schedule.stream:
    pl, _ = stream.plan(program)
    if pl is not None:
      sched = Schedule(**{**asdict(schedule), 'waves_per_block': schedule.stream_waves,
                          'k_block': schedule.stream_k_block})
      emitted = stream.emit(program, context, base_name, sched, pl)
      if emitted.lds_bytes <= LDS_LIMIT:
        return emitted
  if program.epilogue is not None:
    raise GenerationError('pointwise statements fuse only into a streaming chain')
  emitted = _Emitter(program, context, base_name, schedule).emit()
  if emitted.lds_bytes > LDS_LIMIT and schedule.waves_per_block > 1:
    # fewer wavefronts per workgroup: each wavefront owns its LDS slice
    emitted = _Emitter(program, context, base_name,
                       Schedule(**{**asdict(schedule), 'waves_per_block': 1})).emit()
  if emitted.lds_bytes > LDS_LIMIT:
    raise GenerationError(f'tuned emitter: {emitted.lds_bytes} bytes of LDS per workgroup')
  return emitted


# LDS per workgroup on gfx950
LDS_LIMIT = 160 * 1024
