"""Matrix-core forms of the streaming chains (precision study): FP32 MFMA with FP32 accumulation,
and split BF16x3 MFMA with a compensating second accumulator.

Applies to the chains the streaming schedule recognises, with dense shared operands::

    D_e = sum_t U_t · (L_e · W_t,e)          (U_t shared b x b, L_e per element b x q,
                                               W_t,e per element q x q, dense or sparse)

The per-element temporaries ``L_e · W_t,e`` (the star-matrix step) stay in FP64 on the vector
units; the products with the shared matrices run on MFMA. Inputs and outputs keep the FP64 layout
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


def unsupported_reason(program, pl):
  if pl is None:
    return 'not a streaming chain'
  if program.fp_type != FloatingPointType.DOUBLE:
    return 'the matrix-core forms take FP64 chains'
  if any(t.kind != 'fused' for t in pl.terms):
    return 'a term without a per-element temporary'
  if any(t.U.is_sparse() for t in pl.terms):
    return 'sparse shared operand'
  if any(tuple(t.product.k_ranges[0]) != (pl.klo, pl.khi) for t in pl.terms):
    return 'terms with different contraction ranges'
  if pl.L is None or pl.L.arg.addressing != Addressing.STRIDED:
    return 'streamed operand is not strided'
  if pl.l_cols[1] - pl.l_cols[0] != pl.q or pl.q > 16:
    return 'temporary columns'
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
  return None


class _Mfma:
  def __init__(self, program, context, base_name, pl, mode):
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
    self.NKB = (self.K + KB - 1) // KB
    self.kstep = 4 if mode == 'fp32' else 16      # k per MFMA: 16x16x4 f32, 16x16x16 bf16
    self.KS = self.NKB * (KB // self.kstep)
    self.T = len(pl.terms)
    self.prefetch_rows = mode != 'bf16x3'

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
    threads = WAVE * WAVES
    frag = T * KS * RT * WAVE
    if bf16:
      # FP32 in LDS, split into BF16 terms when read: per lane and fragment the 4 consecutive
      # contraction entries of its row (A) or column (B), one 16-byte read
      kbytes = frag * 16
      pitch = 20                          # floats per column of the B tile (16-byte rows)
    else:
      kbytes = frag * 4
      pitch = 17
    tbytes = 64 * pitch * 4
    lds_bytes = kbytes + WAVES * 2 * tbytes
    self.lds_bytes = lds_bytes
    w('typedef float tf_f4 __attribute__((ext_vector_type(4)));')
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
    w(f'// matrix-core form ({self.mode}): {WAVES} wavefronts, {self.G} elements x {self.q} '
      f'columns each, {RT}x{CT} tiles of 16x16, contraction blocks of {KB}')
    w('extern __shared__ __align__(16) unsigned char lds_raw[];')
    if bf16:
      w('float4* __restrict__ kf4 = reinterpret_cast<float4*>(lds_raw);')
    else:
      w('float* __restrict__ kf = reinterpret_cast<float*>(lds_raw);')
    w('const int wave = threadIdx.x / 64;')
    w('const int lane = threadIdx.x % 64;')
    for b in range(2):
      base = f'lds_raw + {kbytes} + (wave * 2 + {b}) * {tbytes}'
      w(f'float* __restrict__ tf{b} = reinterpret_cast<float*>({base});')
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
    w(f'const size_t stride = (size_t)gridDim.x * {WAVES};')
    with w.block('for (size_t grp = (size_t)blockIdx.x * ' + str(WAVES) + ' + wave; grp < ngroups; '
                 'grp += stride)'):
      self.group(pitch)
    w._depth -= 1
    w('}')
    traffic = element_traffic(self.prog)
    sched = {'kind': f'mfma_{self.mode}', 'waves_per_block': WAVES, 'elements_per_wave': G,
             'row_tiles': RT, 'column_tiles': CT, 'k_block': KB, 'lds_bytes': lds_bytes}
    return Emitted(kernel_name=self.kname, kernel_src=w.src(), block=(threads, 1, 1),
                   elements_per_block=G * WAVES, lds_bytes=lds_bytes, emitter='mfma',
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
      if self.mode == 'fp32':
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
    w(f'const double* __restrict__ lq = g_{L.arg.name} + (col + {pl.l_cols[0] - L.origin[1]}) * '
      f'{L.strides[1]};')
    # weights of every slot of every term (FP64)
    for ti, t in enumerate(pl.terms):
      for si, (src, pv) in enumerate(self.tables[ti]):
        w(f'const double w{ti}_{si} = {pv} >= 0 ? (double)g_{t.W.arg.name}[{pv} < 0 ? 0 : {pv}] '
          f': 0.0;')
    w('tf_f4 acc[' + str(RT * CT) + '];')
    if bf16:
      w('tf_f4 acr[' + str(RT * CT) + '];')
    w('#pragma unroll')
    with w.block(f'for (int i = 0; i < {RT * CT}; ++i)'):
      w('acc[i] = tf_f4{0.f, 0.f, 0.f, 0.f};')
      if bf16:
        w('acr[i] = tf_f4{0.f, 0.f, 0.f, 0.f};')
    # software pipeline over the (block, term) steps: the MFMAs of step s read B tile s % 2
    # while the lanes form step s + 1's temporary rows into the other tile; the streamed rows of
    # block kb + 1 are loaded when block kb starts
    steps = [(kb, t) for kb in range(self.NKB) for t in range(self.T)]
    # without prefetch the next block's rows are loaded one step ahead (fewer live registers)
    nxt_loads = lambda si: (si + 1 < len(steps) and steps[si + 1][0] != steps[si][0])
    w(f'double qa[{KB}], qb[{KB}];')
    self.load_rows(0, 'qa')
    self.t_rows(steps[0], 'qa', 0, pitch, range(KB))
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
      per = (len(mf) + KB - 1) // KB
      for j in range(KB):
        if nxt is not None:
          self.t_rows(nxt, qcur if nxt[0] == kb else qnxt, (si + 1) % 2, pitch, [j])
        for line in mf[j * per:(j + 1) * per]:
          w(line)
        # keeps the compiler from hoisting every fragment load and temporary row of the step
        w('__builtin_amdgcn_sched_barrier(0);')
      self.tile_fence()
    self.store()

  def load_rows(self, kb, dst):
    w = self.w
    pl = self.pl
    L = pl.L
    for j in range(KB):
      k = kb * KB + j
      if k < self.K:
        w(f'{dst}[{j}] = lq[{(k + pl.klo - L.origin[0]) * L.strides[0]}];')
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
    tabs = self.tables[ti]
    for j in js:
      v = f'tv{ti}_{j}'
      w(f'{{ const int qlo = __double2loint({q}[{j}]); const int qhi = __double2hiint({q}[{j}]);')
      w(f'  double {v} = 0.0;')
      for si, (src, pv) in enumerate(tabs):
        w(f'  {v} = __builtin_fma(__hiloint2double(__builtin_amdgcn_ds_bpermute({src}, qhi), '
          f'__builtin_amdgcn_ds_bpermute({src}, qlo)), w{ti}_{si}, {v});')
      w(f'  tf{tile}[gcol * {pitch} + {j}] = (float){v}; }}')

  def mfma_list(self, kb, ti, tile, pitch):
    """Source lines of step (kb, ti)'s MFMAs from B tile ``tile``: fragment loads, then the
    MFMAs, as statements of the enclosing scope (names are unique per step)."""
    RT, CT = self.RT, self.CT
    nks = KB // self.kstep
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
          out.append(f'const float bv{ks}_{ct}{sfx} = tf{tile}[({ct * 16} + lane % 16) * {pitch} + '
                     f'{ks * 4} + lane / 16];')
    for ks in range(nks):
      for rt in range(RT):
        frag = f'{((ti * self.KS + kb * nks + ks) * RT + rt) * 64} + lane'
        if self.mode == 'bf16x3':
          a = f'{ks}_{rt}{sfx}'
          out.append(f'tf_s4 ah{a}, am{a}, al{a};')
          out.append(f'tf_split3(kf4[{frag}], ah{a}, am{a}, al{a});')
        else:
          out.append(f'const float av{ks}_{rt}{sfx} = kf[{frag}];')
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
            out.append(f'{a} = __builtin_amdgcn_mfma_f32_16x16x4f32(av{ks}_{rt}{sfx}, '
                       f'bv{ks}_{ct}{sfx}, {a}, 0, 0, 0);')
    return out

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
          w('const int row = rt * 16 + 4 * (lane / 16) + jj;')
          w(f'if (row >= {self.rows}) continue;')
          w(f'const size_t off = (size_t)(row + {pl.rows[0] - o0}) * {s0} + '
            f'(size_t)(co + {pl.cols[0] - o1}) * {s1};')
          if self.mode == 'bf16x3':
            w(f'double v = (double)acc[rt * {CT} + ct][jj] + (double)acr[rt * {CT} + ct][jj];')
          else:
            w(f'double v = (double)acc[rt * {CT} + ct][jj];')
          if pl.init == 'load':
            w('v += dst[off];')
          w('dst[off] = v;')

  def grid_setup(self, lds_bytes, threads):
    return persistent_grid(self.kname, lds_bytes, threads, self.G * WAVES,
                           f'launcher_{self.base_name}')


def emit(program, context, base_name, mode):
  from . import stream
  pl, why = stream.plan(program)
  reason = unsupported_reason(program, pl) if pl is not None else why
  if reason is not None:
    return None, reason
  return _Mfma(program, context, base_name, pl, mode).emit(), None
