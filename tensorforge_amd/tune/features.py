"""Feature vector of (chain, batch, candidate) for the cost model: generation-time facts only
(nothing that needs the kernel compiled or run)."""
import math

import numpy as np

from ..common.basic_types import FloatingPointType

NAMES = [
  'log2_batch', 'fp64', 'products', 'max_rows', 'max_cols', 'sum_k', 'log2_shared_nnz',
  'log2_dense_terms', 'density', 'log2_bytes_per_element', 'log2_flops_per_element',
  'log2_work', 'intensity', 'kind_linear', 'kind_fused', 'kind_three_index',
  'em_plain', 'em_column', 'em_stream', 'waves', 'k_block', 'lookahead', 'log2_lds_bytes',
  'elements_per_block', 'loop_mode', 'log2_fma_sites', 'log2_groups_per_cu',
]


def chain_facts(generator):
  prog = generator.get_program()
  rows = cols = sum_k = 0
  nnz = dense = 0
  for p in prog.products:
    rows = max(rows, p.n_ranges[0][1] - p.n_ranges[0][0])
    if len(p.n_ranges) > 1:
      cols = max(cols, p.n_ranges[1][1] - p.n_ranges[1][0])
    for lo, hi in p.k_ranges:
      sum_k += hi - lo
    for f in p.factors:
      if f.src == 'global' and not f.arg.per_element:
        vol = int(np.prod(f.tensor.shape))
        nnz += len(f.packed) if f.is_sparse() else vol
        dense += vol
  return rows, cols, sum_k, nnz, dense, len(prog.products)


def features(generator, candidate, batch, info):
  em = generator.get_emitted()
  rows, cols, sum_k, nnz, dense, nprod = chain_facts(generator)
  bpe = max(1, em.bytes_per_element)
  flops = max(1, generator.get_program().flops_per_element())
  sched = em.schedule or {}
  kind = info.get('kind', 'linear')
  emitter = 'plain' if candidate.emitter == 'plain' else \
      ('stream' if sched.get('kind') == 'stream' else 'column')
  knobs = candidate.knobs()
  waves = sched.get('waves_per_block', knobs.get('waves_per_block', 1))
  per_block = max(1, em.elements_per_block)
  src = generator.get_kernel() or ''
  sites = src.count('v_fmac_') + src.count('__builtin_fma') + src.count('fma(') + 1
  groups = batch / per_block / 256.0
  return [
    math.log2(batch), float(generator._context.fp_type == FloatingPointType.DOUBLE), nprod,
    rows, cols, sum_k, math.log2(nnz + 1), math.log2(dense + 1), nnz / max(1, dense),
    math.log2(bpe), math.log2(flops), math.log2(flops * batch), flops / bpe,
    float(kind == 'linear'), float(kind == 'fused'), float(kind == 'three_index'),
    float(emitter == 'plain'), float(emitter == 'column'), float(emitter == 'stream'),
    waves, sched.get('k_block', knobs.get('k_block', 0)),
    knobs.get('stream_lds_lookahead', 0) if emitter == 'stream' else 0,
    math.log2(em.lds_bytes + 1), per_block, float(sched.get('loop', False)),
    math.log2(sites), math.log2(groups + 1e-3),
  ]


# knobs of the candidate label that the generation-time facts above do not separate
LABEL_NAMES = ['pair_reads', 'row_loads', 'lds_unpadded', 'k16_blocks']


def label_features(label):
  """Indicator features of a candidate's label (``space.candidate_space`` names)."""
  parts = label.split('_')
  return [float('pair' in parts), float('rows' in parts), float('nopad' in parts),
          float(label.startswith('stream') and 'k16' in parts)]


# occupancy and load balance, derived from the generation-time facts above (a workgroup's LDS
# bytes, wavefronts and elements, the batch) for an MI355X of 256 CUs with 160 KB of LDS each;
# the second set also bounds occupancy by an estimate of the registers a lane holds
DERIVED_NAMES = ['workgroups_per_cu', 'waves_per_cu', 'log2_rounds', 'last_round_fill',
                 'log2_vgpr_estimate', 'waves_per_simd_by_vgpr', 'workgroups_per_cu_by_vgpr',
                 'log2_rounds_by_vgpr', 'last_round_fill_by_vgpr']

CUS = 256
LDS_PER_CU = 160 * 1024


def derived_features(v):
  """Occupancy and load-balance estimates from feature vector ``v`` (``NAMES`` order): resident
  workgroups per CU by LDS (at most 8), resident wavefronts per CU (at most 32), the number of
  rounds of workgroups the batch needs on the whole GPU, and how full the last round is; then
  the same with occupancy also bounded by registers (the estimate, wavefronts per SIMD,
  workgroups per CU, whole rounds of the persistent grid, fill of the last round)."""
  if len(v) != len(NAMES):       # not a generated feature vector (hand-made records in tests)
    return [0.0] * len(DERIVED_NAMES)
  f = dict(zip(NAMES, v))
  lds = 2.0 ** f['log2_lds_bytes'] - 1.0
  waves = max(1.0, f['waves'])
  wg = float(max(1, min(8, int(LDS_PER_CU // lds) if lds >= 1 else 8, int(32 // waves))))
  batch = 2.0 ** f['log2_batch']
  groups = batch / max(1.0, f['elements_per_block'])
  rounds = groups / (CUS * wg)
  full = rounds / math.ceil(rounds) if rounds > 0 else 1.0
  # registers: the column-per-lane and streaming forms keep the lane's destination column
  # (rows x 2 words in FP64) plus a fixed share for streamed rows, addresses and temporaries;
  # 512 per SIMD lane give the resident wavefronts per SIMD (at most 8), four SIMDs per CU
  words = 2 if f['fp64'] else 1
  vg = 64.0 if f['em_plain'] else f['max_rows'] * words + (40 if f['em_stream'] else 24)
  wps = max(1, min(8, int(512 // max(vg, 24.0))))
  wg_v = max(1, min(int(LDS_PER_CU // lds) if lds >= 1 else 32,
                    int(wps * 4 // waves) if waves <= wps * 4 else 1))
  grid = CUS * wg_v
  rounds_v = math.ceil(groups / grid) if groups > 0 else 1
  fill_v = groups / (rounds_v * grid)
  return [wg, wg * waves, math.log2(rounds + 1e-3), full,
          math.log2(vg), float(wps), float(wg_v), math.log2(rounds_v), fill_v]
