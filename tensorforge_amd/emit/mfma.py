"""Matrix-core forms of the streaming chains: FP64 MFMA (``v_mfma_f64_16x16x4f64``, the same
arithmetic type as the vector form, for shared operands dense enough that their padded 16x16
tiles cost less than one DPP FMA per stored entry), and for the precision study FP32 MFMA with
FP32 accumulation and split BF16x3 MFMA with a compensating second accumulator.

Applies to the chains the streaming schedule recognises, with dense shared operands::

    D_e = sum_t U_t · (L_e · W_t,e)          (U_t shared b x b, L_e per element b x q,
                                               W_t,e per element q x q, dense or sparse)
    D_e = [D_e +] U · B_e                    (one product with a per-element right operand)

The per-element temporaries ``L_e · W_t,e`` (the star-matrix step) stay in FP64 on the vector
units; the products with the shared matrices run on MFMA. In FP64 the shared operands sit in LDS
as doubles, so the contraction blocks shrink to 8 rows when 16 would not fit. Inputs and outputs keep the FP64 layout
and launcher of the chain, so the launcher is a drop-in for the FP64 one.

Mapping: a workgroup of 4 wavefronts stays resident; each wavefront owns ``64 // q`` elements, i.e.
``G·q`` output columns = ``ceil(G·q / 16)`` MFMA column tiles, and all row tiles of them. The
shared matrices are converted once per workgroup into LDS in MFMA A-operand order. Per block of 16
contraction rows and per term, each lane forms its column of the temporary (FP64, operands of other
columns through ``ds_bpermute``), rounds or splits it into a per-wavefront LDS tile in B-operand
order, and the wavefront issues the MFMAs of that block.

Split BF16x3 (three-term split, six products; the scheme of the reference's unused design input
``backend/instructions/compute/primitives/nvidia.py:126-170``): both operands are kept in FP32 in
LDS and split into three BF16 terms when a fragment is read, ``x = hi + mid + lo`` with
``hi = rne_bf16(x)``, ``mid = rne_bf16(x - hi)``, ``lo = x - hi - mid``. Both differences are exact
in FP32, and ``lo`` has at most 8 significant bits, so the three terms hold the FP32 value exactly.
``a·b ~ hi_a·hi_b + (hi_a·mid_b + mid_a·hi_b + hi_a·lo_b + lo_a·hi_b + mid_a·mid_b)``: the dropped
``mid·lo``, ``lo·mid`` and ``lo·lo`` products are below 2^-24 of the leading one. The first product
accumulates in one FP32 accumulator and the five corrections in a second. The two are added in FP64
at the end, so the small corrections are not absorbed by the rounding of the large sum.
"""
from ..common.basic_types import Addressing, FloatingPointType
from . import ir
from .launcher import Emitted, Writer, kernel_params, persistent_grid
from .traffic import element_traffic

WAVE = 64
WAVES = 4
KB = 16            # contraction rows per block


LDS_LIMIT = 160 * 1024


def _shape(pl, mode, sched=None):
  """(wavefronts per workgroup, contraction rows per block). The FP32 and BF16x3 forms: 4 and 16.
  FP64: four wavefronts (one per SIMD: the registers hold the accumulators, the destination's
  old values loaded at the start of each group and the streamed rows of three blocks), with
  16-row blocks, or 8 when the shared operands as doubles and 16-row B tiles do not fit in LDS.
  ``sched.mfma_waves`` and ``sched.mfma_k_block`` override the choice."""
  if mode != 'fp64':
    return WAVES, KB
  if sched is not None and sched.mfma_waves and sched.mfma_k_block:
    return sched.mfma_waves, sched.mfma_k_block
  shapes = ((4, 16), (4, 8))
  for waves, kb in shapes:
    if _lds_bytes(pl, mode, kb, waves) <= LDS_LIMIT:
      return waves, kb
  return WAVES, 8


def _lds_bytes(pl, mode, kb, waves=WAVES):
  """LDS of the form: the shared operands in A-operand order plus two B tiles per wavefront."""
  K = pl.khi - pl.klo
  ks = (K + kb - 1) // kb * (kb // 4 if mode != 'bf16x3' else kb // 16)
  rt = (pl.rows[1] - pl.rows[0] + 15) // 16
  frag = len(pl.terms) * ks * rt * WAVE
  if mode == 'bf16x3':
    return frag * 16 + waves * 2 * 64 * 20 * 4
  if mode == 'fp64':
    return frag * 8 + waves * 2 * 64 * (kb + 1) * 8
  return frag * 4 + waves * 2 * 64 * 17 * 4


def unsupported_reason(program, pl, mode='fp32', sched=None):
  if pl is None:
    return 'not a streaming chain'
  if program.fp_type != FloatingPointType.DOUBLE:
    return 'the matrix-core forms take FP64 chains'
  kinds = {t.kind for t in pl.terms}
  if kinds == {'global'}:
    if len(pl.terms) != 1:
      return 'more than one product with a per-element right operand'
  elif kinds != {'fused'}:
    return 'a term without a per-element temporary'
  if any(t.U.is_sparse() for t in pl.terms):
    return 'sparse shared operand'
  if any(tuple(t.product.k_ranges[0]) != (pl.klo, pl.khi) for t in pl.terms):
    return 'terms with different contraction ranges'
  if kinds == {'fused'}:
    if pl.L is None or pl.L.arg.addressing != Addressing.STRIDED:
      return 'streamed operand is not strided'
    if pl.l_cols[1] - pl.l_cols[0] != pl.q:
      return 'temporary columns'
  if pl.q > 16:
    return 'more than 16 columns'
  if pl.rows[1] - pl.rows[0] > 128 or pl.khi - pl.klo > 128:
    return 'more than 128 rows'
  if pl.init == 'load' and pl.beta != 1.0:
    return 'scaled accumulation'
  # register stages after the streamed product (the flux integrals' U_n ... U_2 chain) and
  # their final accumulation are not part of this form
  if pl.post or pl.final_init != 'zero':
    return 'register stages after the streamed product'
  # every per-element operand (the destination and the weights included) is addressed as
  # element * volume + extraOffset
  if pl.dest.addressing != Addressing.STRIDED:
    return 'destination is not strided'
  if any(a.per_element and a.addressing != Addressing.STRIDED for a in program.args):
    return 'a per-element operand is not strided'
  if program.epilogue is not None:
    return 'pointwise statements fused into the chain'
  waves, kb = _shape(pl, mode, sched)
  if waves not in (4, 8) or kb not in ((8, 16) if mode == 'fp64' else (KB,)):
    return f'{waves} wavefronts with {kb}-row blocks'
  if _lds_bytes(pl, mode, kb, waves) > LDS_LIMIT:
    return 'shared operands do not fit in LDS'
  if sched is not None and sched.mfma_dest not in ('', 'acc', 'regs', 'end') or \
      (mode != 'fp64' and sched is not None and sched.mfma_dest not in ('', 'end')):
    return f'destination mode {sched.mfma_dest!r}'
  return None


class _Mfma:
  def __init__(self, program, context, base_name, pl, mode, sched=None):
    self.prog, self.ctx, self.pl, self.mode = program, context, pl, mode
    self.base_name = base_name
    self.kname = f'kernel_{base_name}'
    self.w = Writer()
    self.q = pl.q
    self.G = WAVE // pl.q
    self.cols = self.G * pl.q
    self.CT = (self.cols + 15) // 16
    self.rows = pl.rows[1] - pl.rows[0]
    self.RT = (self.rows + 15) // 16
    self.K = pl.khi - pl.klo
    self.waves, self.KB = _shape(pl, mode, sched)
    self.NKB = (self.K + self.KB - 1) // self.KB
    # k per MFMA: 16x16x4 f64 and f32, 16x16x16 bf16
    self.kstep = 16 if mode == 'bf16x3' else 4
    self.KS = self.NKB * (self.KB // self.kstep)
    self.T = len(pl.terms)
    self.prefetch_rows = mode != 'bf16x3'
    # one product with a per-element right operand: its rows go to the B tile as they are
    self.B = pl.terms[0].B if pl.terms[0].kind == 'global' else None
    self.real = 'double' if mode == 'fp64' else 'float'
    # where an accumulating product's old destination values wait: 'acc' (FP64 default: the
    # accumulators start from them, loaded at the start of the group), 'regs' (registers of their
    # own, loaded after the group's first streamed block, added at the end), 'end' (loaded and
    # added at the end; the FP32 and BF16x3 forms)
    dest = (sched.mfma_dest if sched is not None else '') or ('regs' if mode == 'fp64' else 'end')
    self.dest_mode = dest if pl.init == 'load' else 'end'
    self.preload = self.dest_mode == 'acc'

  # --------------------------------------------------------------- per-lane star slots
  def slot_tables(self, t):
    """[(m per column, position per column or -1)] for term t's weights."""
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
        pos = t.W.packed.get(tuple(idx)) if t.W.is_sparse() else t.W.offset(idx)
        if pos is not None:
          ms.append((m, pos))
      cols.append(ms)
    n = max(len(c) for c in cols)
    out = []
    for s in range(n):
      m = [c[s][0] if s < len(c) else (c[0][0] if c else mlo) for c in cols]
      p = [c[s][1] if s < len(c) else -1 for c in cols]
      out.append((m, p))
    return out

  @staticmethod
  def select(values):
    if all(v == values[0] for v in values):
      return str(values[0])
    expr = str(values[-1])
    for c in range(len(values) - 2, -1, -1):
      expr = f'(col == {c} ? {values[c]} : {expr})'
    return expr

  # --------------------------------------------------------------- emission
  def emit(self):
    w = self.w
    pl = self.pl
    T, KS, RT, CT = self.T, self.KS, self.RT, self.CT
    bf16 = self.mode == 'bf16x3'
    fp64 = self.mode == 'fp64'
    threads = WAVE * self.waves
    frag = T * KS * RT * WAVE
    if bf16:
      # FP32 in LDS, split into BF16 terms when read: per lane and fragment the 4 consecutive
      # contraction entries of its row (A) or column (B), one 16-byte read
      kbytes = frag * 16
      pitch = 20                          # floats per column of the B tile (16-byte rows)
      tbytes = 64 * pitch * 4
    elif fp64:
      kbytes = frag * 8
      pitch = self.KB + 1                 # doubles per column of the B tile
      tbytes = 64 * pitch * 8
    else:
      kbytes = frag * 4
      pitch = 17
      tbytes = 64 * pitch * 4
    lds_bytes = kbytes + self.waves * 2 * tbytes
    assert lds_bytes == _lds_bytes(pl, self.mode, self.KB, self.waves)
    self.lds_bytes = lds_bytes
    w('typedef float tf_f4 __attribute__((ext_vector_type(4)));')
    if fp64:
      w('typedef double tf_d4 __attribute__((ext_vector_type(4)));')
    w('typedef short tf_s4 __attribute__((ext_vector_type(4)));')
    w('#ifndef TF_BF16_SPLIT3')
    w('#define TF_BF16_SPLIT3')
    # tf_rne_top: an FP32 value rounded to BF16 (nearest, ties to even), kept in the top 16 bits
    # (finite inputs). tf_split3: 4 FP32 values split into BF16 hi / mid / lo fragments (entry j
    # in halfword j)
    w('__device__ __forceinline__ unsigned tf_rne_top(float x) {')
    w('  const unsigned u = __float_as_uint(x);')
    w('  return (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;')
    w('}')
    w('__device__ __forceinline__ void tf_split3(const float4 x, tf_s4& h, tf_s4& m, tf_s4& l) {')
    w('  const float v[4] = {x.x, x.y, x.z, x.w};')
    w('  unsigned a[4], b[4], c[4];')
    w('#pragma unroll')
    w('  for (int j = 0; j < 4; ++j) {')
    w('    a[j] = tf_rne_top(v[j]);')
    w('    const float r = v[j] - __uint_as_float(a[j]);')
    w('    b[j] = tf_rne_top(r);')
    w('    c[j] = __float_as_uint(r - __uint_as_float(b[j]));')
    w('  }')
    w('  h = __builtin_bit_cast(tf_s4, make_uint2((a[0] >> 16) | (a[1] & 0xffff0000u), '
      '(a[2] >> 16) | (a[3] & 0xffff0000u)));')
    w('  m = __builtin_bit_cast(tf_s4, make_uint2((b[0] >> 16) | (b[1] & 0xffff0000u), '
      '(b[2] >> 16) | (b[3] & 0xffff0000u)));')
    w('  l = __builtin_bit_cast(tf_s4, make_uint2((c[0] >> 16) | (c[1] & 0xffff0000u), '
      '(c[2] >> 16) | (c[3] & 0xffff0000u)));')
    w('}')
    w('#endif')
    params = ', '.join(kernel_params(self.prog, self.ctx))
    w(f'extern "C" __global__ void __launch_bounds__({threads}, 1) {self.kname}({params})')
    w('{')
    w._depth += 1
    w(f'// matrix-core form ({self.mode}): {self.waves} wavefronts, {self.G} elements x {self.q} '
      f'columns each, {RT}x{CT} tiles of 16x16, contraction blocks of {self.KB}')
    w('extern __shared__ __align__(16) unsigned char lds_raw[];')
    if bf16:
      w('float4* __restrict__ kf4 = reinterpret_cast<float4*>(lds_raw);')
    else:
      w(f'{self.real}* __restrict__ kf = reinterpret_cast<{self.real}*>(lds_raw);')
    w('const int wave = threadIdx.x / 64;')
    w('const int lane = threadIdx.x % 64;')
    for b in range(2):
      base = f'lds_raw + {kbytes} + (wave * 2 + {b}) * {tbytes}'
      w(f'{self.real}* __restrict__ tf{b} = reinterpret_cast<{self.real}*>({base});')
    self.stage_shared(frag, threads)
    w('__syncthreads();')
    q, G = self.q, self.G
    w(f'const int slot = lane / {q} < {G} ? lane / {q} : {G - 1};')
    w(f'const int col = lane / {q} < {G} ? lane - slot * {q} : 0;')
    w(f'const int gcol = slot * {q} + col;')
    # bpermute source lanes of every slot table
    self.tables = []
    src_names = {}
    for ti, t in enumerate(pl.terms):
      if t.kind != 'fused':
        self.tables.append([])
        continue
      tabs = []
      for si, (m, p) in enumerate(self.slot_tables(t)):
        key = tuple(m)
        if key not in src_names:
          v = f'src{len(src_names)}'
          w(f'const int {v} = (slot * {q} + ({self.select(list(m))} - {pl.l_cols[0]})) * 4;')
          src_names[key] = v
        pv = f'wp{ti}_{si}'
        w(f'const int {pv} = {self.select(p)};')
        tabs.append((src_names[key], pv))
      self.tables.append(tabs)
    w(f'const size_t ngroups = (numElements + {G - 1}) / {G};')
    w(f'const size_t stride = (size_t)gridDim.x * {self.waves};')
    with w.block('for (size_t grp = (size_t)blockIdx.x * ' + str(self.waves) + ' + wave; grp < ngroups; '
                 'grp += stride)'):
      self.group(pitch)
    w._depth -= 1
    w('}')
    traffic = element_traffic(self.prog)
    sched = {'kind': f'mfma_{self.mode}', 'waves_per_block': self.waves, 'elements_per_wave': G,
             'row_tiles': RT, 'column_tiles': CT, 'k_block': self.KB, 'lds_bytes': lds_bytes}
    return Emitted(kernel_name=self.kname, kernel_src=w.src(), block=(threads, 1, 1),
                   elements_per_block=G * self.waves, lds_bytes=lds_bytes, emitter='mfma',
                   schedule=sched, bytes_per_element=traffic.per_element,
                   shared_bytes=traffic.shared, grid_override='tf_grid',
                   extra_launch_setup=self.grid_setup(lds_bytes, threads))

  def shared_value(self, t_expr, row, k):
    """FP64 value of U_t(row, k) (relative indices), 0 outside the ranges."""
    pl = self.pl
    expr = '0.0'
    for ti in range(self.T - 1, -1, -1):
      t = pl.terms[ti]
      U = t.U
      nd, kd = t.u_dims
      pos = (f'({row} + {pl.rows[0] - U.origin[nd]}) * {U.strides[nd]} + ({k} + '
             f'{pl.klo - U.origin[kd]}) * {U.strides[kd]}')
      expr = f'({t_expr} == {ti} ? (double){U.arg.name}[{pos}] : {expr})'
    return f'(({row}) < {self.rows} && ({k}) < {self.K} ? {expr} : 0.0)'

  def stage_shared(self, frag, threads):
    w = self.w
    RT, KS = self.RT, self.KS
    with w.block(f'for (int e = threadIdx.x; e < {frag}; e += {threads})'):
      w('const int l = e % 64;')
      w(f'const int rt = (e / 64) % {RT};')
      w(f'const int ks = (e / {64 * RT}) % {KS};')
      w(f'const int t = e / {64 * RT * KS};')
      w('const int row = rt * 16 + l % 16;')
      if self.mode == 'fp64':
        w(f'kf[e] = {self.shared_value("t", "row", "ks * 4 + l / 16")};')
      elif self.mode == 'fp32':
        w(f'kf[e] = (float){self.shared_value("t", "row", "ks * 4 + l / 16")};')
      else:
        w('float x[4];')
        w('#pragma unroll')
        with w.block('for (int j = 0; j < 4; ++j)'):
          w(f'x[j] = (float){self.shared_value("t", "row", "ks * 16 + 4 * (l / 16) + j")};')
        w('kf4[e] = make_float4(x[0], x[1], x[2], x[3]);')

  def group(self, pitch):
    w = self.w
    pl = self.pl
    q, G, RT, CT = self.q, self.G, self.RT, self.CT
    bf16 = self.mode == 'bf16x3'
    L = pl.L
    w(f'const size_t e0 = grp * {G};')
    w('const size_t batchId = e0 + slot;')
    w(f'const size_t eSafe = batchId < numElements ? batchId : 0;')
    # per-element operands of this lane's element
    for a in self.prog.args:
      if a.kind != ir.SRC_GLOBAL or not a.per_element:
        continue
      ty = f'{"const " if a.const else ""}{a.dtype}'
      w(f'{ty}* __restrict__ g_{a.name} = {a.name} + eSafe * {a.element_stride}ull + '
        f'{a.name}_extraOffset;')
    if self.mode == 'fp64':
      w(f'const double* __restrict__ lq = {self.column_ptr("eSafe")};')
    else:
      w(f'const double* __restrict__ lq = g_{L.arg.name} + (col + {pl.l_cols[0] - L.origin[1]}) '
        f'* {L.strides[1]};')
    # weights of every slot of every term (FP64)
    for ti, t in enumerate(pl.terms):
      for si, (src, pv) in enumerate(self.tables[ti]):
        w(f'const double w{ti}_{si} = {pv} >= 0 ? (double)g_{t.W.arg.name}[{pv} < 0 ? 0 : {pv}] '
          f': 0.0;')
    acc_t = 'tf_d4' if self.mode == 'fp64' else 'tf_f4'
    w(f'{acc_t} acc[{RT * CT}];')
    if bf16:
      w('tf_f4 acr[' + str(RT * CT) + '];')
    w('#pragma unroll')
    with w.block(f'for (int i = 0; i < {RT * CT}; ++i)'):
      w(f'acc[i] = {acc_t}{{0.0, 0.0, 0.0, 0.0}};' if self.mode == 'fp64' else
        'acc[i] = tf_f4{0.f, 0.f, 0.f, 0.f};')
      if bf16:
        w('acr[i] = tf_f4{0.f, 0.f, 0.f, 0.f};')
    if self.preload:
      self.load_destination()
    # software pipeline over the (block, term) steps: the MFMAs of step s read B tile s % 2
    # while the lanes form step s + 1's temporary rows into the other tile; the streamed rows of
    # block kb + 1 are loaded when block kb starts
    steps = [(kb, t) for kb in range(self.NKB) for t in range(self.T)]
    # without prefetch the next block's rows are loaded one step ahead (fewer live registers)
    nxt_loads = lambda si: (si + 1 < len(steps) and steps[si + 1][0] != steps[si][0])
    KBr = self.KB
    w(f'double qa[{KBr}], qb[{KBr}];')
    self.load_rows(0, 'qa')
    if self.dest_mode == 'regs':
      self.load_destination('cpre')
    self.t_rows(steps[0], 'qa', 0, pitch, range(KBr))
    self.tile_fence()
    for si, (kb, t) in enumerate(steps):
      qcur = 'qa' if kb % 2 == 0 else 'qb'
      qnxt = 'qb' if kb % 2 == 0 else 'qa'
      w(f'// step {si}: block {kb}, term {t}')
      if t == 0 and kb + 1 < self.NKB and self.prefetch_rows:
        self.load_rows(kb + 1, qnxt)
      if not self.prefetch_rows and nxt_loads(si):
        self.load_rows(kb + 1, qnxt)
      nxt = steps[si + 1] if si + 1 < len(steps) else None
      mf = self.mfma_list(kb, t, si % 2, pitch)
      per = (len(mf) + KBr - 1) // KBr
      for j in range(KBr):
        if nxt is not None:
          self.t_rows(nxt, qcur if nxt[0] == kb else qnxt, (si + 1) % 2, pitch, [j])
        for line in mf[j * per:(j + 1) * per]:
          w(line)
        # keeps the compiler from hoisting every fragment load and temporary row of the step
        w('__builtin_amdgcn_sched_barrier(0);')
      self.tile_fence()
    self.store()

  def column_ptr(self, esafe):
    """The lane's column of the streamed per-element operand (the right operand of a GEMM, or
    the operand the temporaries are formed from) for element ``esafe``."""
    pl = self.pl
    if self.B is not None:
      X, nd = self.B, pl.terms[0].b_dims[1]
      c0 = pl.cols[0]
    else:
      X, nd = pl.L, 1
      c0 = pl.l_cols[0]
    a = X.arg
    return (f'{a.name} + {esafe} * {a.element_stride}ull + {a.name}_extraOffset + '
            f'(col + {c0 - X.origin[nd]}) * {X.strides[nd]}')

  def load_rows(self, kb, dst):
    """Rows of contraction block ``kb`` of the lane's column (at ``lq``) into ``dst``."""
    w = self.w
    pl = self.pl
    if self.B is not None:
      kd = pl.terms[0].b_dims[0]
      origin, stride = self.B.origin[kd], self.B.strides[kd]
    else:
      origin, stride = pl.L.origin[0], pl.L.strides[0]
    for j in range(self.KB):
      k = kb * self.KB + j
      if k < self.K:
        w(f'{dst}[{j}] = lq[{(k + pl.klo - origin) * stride}];')
      else:
        w(f'{dst}[{j}] = 0.0;')

  def tile_fence(self):
    w = self.w
    w('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");')
    w('__builtin_amdgcn_wave_barrier();')
    w('__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')

  def t_rows(self, step, q, tile, pitch, js):
    """Rows ``js`` of step (kb, t)'s temporary: FP64, from the lanes' rows via bpermute, rounded
    or split into B tile ``tile``."""
    w = self.w
    kb, ti = step
    cast = '' if self.mode == 'fp64' else '(float)'
    if self.B is not None:
      for j in js:
        w(f'tf{tile}[gcol * {pitch} + {j}] = {cast}{q}[{j}];')
      return
    tabs = self.tables[ti]
    for j in js:
      v = f'tv{ti}_{j}'
      w(f'{{ const int qlo = __double2loint({q}[{j}]); const int qhi = __double2hiint({q}[{j}]);')
      w(f'  double {v} = 0.0;')
      for si, (src, pv) in enumerate(tabs):
        w(f'  {v} = __builtin_fma(__hiloint2double(__builtin_amdgcn_ds_bpermute({src}, qhi), '
          f'__builtin_amdgcn_ds_bpermute({src}, qlo)), w{ti}_{si}, {v});')
      w(f'  tf{tile}[gcol * {pitch} + {j}] = {cast}{v}; }}')

  def mfma_list(self, kb, ti, tile, pitch):
    """Source lines of step (kb, ti)'s MFMAs from B tile ``tile``: fragment loads, then the
    MFMAs, as statements of the enclosing scope (names are unique per step)."""
    RT, CT = self.RT, self.CT
    nks = self.KB // self.kstep
    out = []
    sfx = f'_{kb}_{ti}'
    for ks in range(nks):
      for ct in range(CT):
        if self.mode == 'bf16x3':
          o = f'(({ct * 16} + lane % 16) * {pitch} + 4 * (lane / 16))'
          b = f'{ks}_{ct}{sfx}'
          out.append(f'tf_s4 bh{b}, bm{b}, bl{b};')
          out.append(f'tf_split3(*reinterpret_cast<const float4*>(tf{tile} + {o}), bh{b}, bm{b}, '
                     f'bl{b});')
        else:
          out.append(f'const {self.real} bv{ks}_{ct}{sfx} = tf{tile}[({ct * 16} + lane % 16) * '
                     f'{pitch} + {ks * 4} + lane / 16];')
    for ks in range(nks):
      for rt in range(RT):
        frag = f'{((ti * self.KS + kb * nks + ks) * RT + rt) * 64} + lane'
        if self.mode == 'bf16x3':
          a = f'{ks}_{rt}{sfx}'
          out.append(f'tf_s4 ah{a}, am{a}, al{a};')
          out.append(f'tf_split3(kf4[{frag}], ah{a}, am{a}, al{a});')
        else:
          out.append(f'const {self.real} av{ks}_{rt}{sfx} = kf[{frag}];')
        for ct in range(CT):
          a = f'acc[{rt * CT + ct}]'
          if self.mode == 'bf16x3':
            r = f'acr[{rt * CT + ct}]'
            ai, bi = f'{ks}_{rt}{sfx}', f'{ks}_{ct}{sfx}'
            mf = '__builtin_amdgcn_mfma_f32_16x16x16bf16_1k'
            out.append(f'{a} = {mf}(ah{ai}, bh{bi}, {a}, 0, 0, 0);')
            # corrections, smallest first: mid·mid, lo·hi, hi·lo, mid·hi, hi·mid
            for x, y in (('am', 'bm'), ('al', 'bh'), ('ah', 'bl'), ('am', 'bh'), ('ah', 'bm')):
              out.append(f'{r} = {mf}({x}{ai}, {y}{bi}, {r}, 0, 0, 0);')
          else:
            mf = '__builtin_amdgcn_mfma_f64_16x16x4f64' if self.mode == 'fp64' else \
                '__builtin_amdgcn_mfma_f32_16x16x4f32'
            out.append(f'{a} = {mf}(av{ks}_{rt}{sfx}, bv{ks}_{ct}{sfx}, {a}, 0, 0, 0);')
    return out

  def dest_tile_loop(self, body):
    """Loops over the lane's accumulator entries (tile, entry jj) that hold a destination entry
    of an element in the batch; ``body(d)`` writes the statements for entry ``acc[i][jj]``
    at ``d`` + ``off``."""
    w = self.w
    pl = self.pl
    q, G, RT, CT = self.q, self.G, self.RT, self.CT
    dp = pl.dest_product
    o0, o1 = dp.dest_origin
    s0, s1 = dp.dest_strides
    d = pl.dest
    w('#pragma unroll')
    with w.block(f'for (int rt = 0; rt < {RT}; ++rt)'):
      w('#pragma unroll')
      with w.block(f'for (int ct = 0; ct < {CT}; ++ct)'):
        w('const int g = ct * 16 + lane % 16;')
        w(f'const int so = g / {q};')
        w(f'const int co = g % {q};')
        w('const size_t eo = e0 + so;')
        w(f'if (g >= {G * q} || eo >= numElements) continue;')
        body(d, o0, o1, s0, s1)

  def load_destination(self, into='acc'):
    """FP64 accumulation into the destination: its old values, loaded before the contraction so
    that their latency overlaps it, into the accumulators or into registers ``into`` of their own
    (the MFMA D layout: entry jj of lane l is row 4 jj + l / 16 of the tile, column l % 16)."""
    w = self.w
    pl = self.pl
    if into != 'acc':
      w(f'tf_d4 {into}[{self.RT * self.CT}];')
      w('#pragma unroll')
      with w.block(f'for (int i = 0; i < {self.RT * self.CT}; ++i)'):
        w(f'{into}[i] = tf_d4{{0.0, 0.0, 0.0, 0.0}};')

    def body(d, o0, o1, s0, s1):
      w(f'const {d.dtype}* __restrict__ src = {d.name} + eo * {d.element_stride}ull + '
        f'{d.name}_extraOffset;')
      w('#pragma unroll')
      with w.block('for (int jj = 0; jj < 4; ++jj)'):
        w('const int row = rt * 16 + 4 * jj + lane / 16;')
        w(f'if (row >= {self.rows}) continue;')
        w(f'{into}[rt * {self.CT} + ct][jj] = src[(size_t)(row + {pl.rows[0] - o0}) * {s0} + '
          f'(size_t)(co + {pl.cols[0] - o1}) * {s1}];')
    self.dest_tile_loop(body)

  def store(self):
    w = self.w
    pl = self.pl
    q, G, RT, CT = self.q, self.G, self.RT, self.CT
    dp = pl.dest_product
    o0, o1 = dp.dest_origin
    s0, s1 = dp.dest_strides
    d = pl.dest
    w('#pragma unroll')
    with w.block(f'for (int rt = 0; rt < {RT}; ++rt)'):
      w('#pragma unroll')
      with w.block(f'for (int ct = 0; ct < {CT}; ++ct)'):
        w('const int g = ct * 16 + lane % 16;')
        w(f'const int so = g / {q};')
        w(f'const int co = g % {q};')
        w('const size_t eo = e0 + so;')
        w(f'if (g >= {G * q} || eo >= numElements) continue;')
        w('if (flags != nullptr && flags[eo] == 0u) continue;')
        w(f'{d.dtype}* __restrict__ dst = {d.name} + eo * {d.element_stride}ull + '
          f'{d.name}_extraOffset;')
        w('#pragma unroll')
        with w.block('for (int jj = 0; jj < 4; ++jj)'):
          # accumulator entry jj of the lane: rows 4 (lane / 16) + jj of the tile for the FP32
          # instruction (four consecutive rows per lane), rows 4 jj + lane / 16 for the FP64 one
          if self.mode == 'fp64':
            w('const int row = rt * 16 + 4 * jj + lane / 16;')
          else:
            w('const int row = rt * 16 + 4 * (lane / 16) + jj;')
          w(f'if (row >= {self.rows}) continue;')
          w(f'const size_t off = (size_t)(row + {pl.rows[0] - o0}) * {s0} + '
            f'(size_t)(co + {pl.cols[0] - o1}) * {s1};')
          if self.mode == 'bf16x3':
            w(f'double v = (double)acc[rt * {CT} + ct][jj] + (double)acr[rt * {CT} + ct][jj];')
          else:
            w(f'double v = (double)acc[rt * {CT} + ct][jj];')
          if self.dest_mode == 'end' and pl.init == 'load':
            w('v += dst[off];')
          elif self.dest_mode == 'regs':
            w(f'v += cpre[rt * {CT} + ct][jj];')
          w('dst[off] = v;')

  def grid_setup(self, lds_bytes, threads):
    return persistent_grid(self.kname, lds_bytes, threads, self.G * self.waves,
                           f'launcher_{self.base_name}')


def emit(program, context, base_name, mode, sched=None):
  from . import stream
  pl, why = stream.plan(program)
  reason = unsupported_reason(program, pl, mode, sched) if pl is not None else why
  if reason is not None:
    return None, reason
  return _Mfma(program, context, base_name, pl, mode, sched).emit(), None
