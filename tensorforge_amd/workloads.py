"""Description lists and seeded inputs of the benchmark workloads.

* ``gemm``: ``C_e += A · B_e``, A b×b shared, B_e and C_e b×q per element (single batched GEMM).
* ``pointwise_update``: ``Q_e += dt · I_e`` entry by entry (the ADER update of the degrees of
  freedom with the time-integrated ones), ``dt`` a by-value scalar.
* ``pointwise_plastic``: a nodal Drucker-Prager return map on the six stress columns of a b×9
  nodal array: mean stress ``m``, deviatoric invariant ``tau = sqrt(J2)``, yield stress
  ``c - m·tanphi``; nodes with ``tau`` above it have their deviator scaled back to the yield
  surface.
* ``ader_volume``: ``Q'_e = Σ_d K_d · Q_e · S_{e,d}``, d = 1..3: K_d b×b shared (sparse in CSR
  value order, or dense), Q_e b×q, S_{e,d} a q×q per-element star matrix with a fixed 24-entry
  pattern. This is the volume chain of SeisSol's ADER-DG step as written in the reference's
  benchmark program ``benchmarks/fused/programs/s3time.cf:43``.

Inputs come from ``numpy.random.default_rng(seed)`` (PCG64) in a fixed draw order, column-major
per element, elements contiguous.
"""
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .common.aux import generate_tmp_tensor
from .common.basic_types import Addressing, FloatingPointType
from .common.matrix.boundingbox import BoundingBox
from .common.matrix.spp import ListSPP
from .common.matrix.tensor import SubTensor, Tensor
from .generators import optree as ot
from .generators.descriptions import ElementwiseDescr, MultilinearDescr

# Elastic-wave star matrix pattern (quantities σxx σyy σzz σxy σyz σxz u v w = 0..8): the
# stress rows couple to velocities, the velocity rows to stresses.
STAR_PATTERN = ([(r, c) for r in range(3) for c in range(6, 9)] +
                [(3, 6), (3, 7), (4, 7), (4, 8), (5, 6), (5, 8)] +
                [(6, 0), (6, 3), (6, 5), (7, 3), (7, 1), (7, 4), (8, 5), (8, 4), (8, 2)])

BASIS_SIZES = {1: 1, 2: 4, 3: 10, 4: 20, 5: 35, 6: 56, 7: 84, 8: 120}


def star_entries():
  """Star pattern entries in column-major order (the packed value order)."""
  return sorted(STAR_PATTERN, key=lambda rc: (rc[1], rc[0]))


@dataclass
class Workload:
  name: str
  descrs: List
  fp_type: FloatingPointType
  tensors: Dict[str, Tensor]
  # names of the launcher operands in launcher order are known after Generator naming
  meta: Dict = field(default_factory=dict)


def gemm(fp_type=FloatingPointType.DOUBLE, b=56, q=9) -> Workload:
  A = Tensor([b, b], Addressing.NONE, alias='A')
  B = Tensor([b, q], Addressing.STRIDED, alias='B')
  C = Tensor([b, q], Addressing.STRIDED, alias='C')
  d = MultilinearDescr(SubTensor(C), [SubTensor(A), SubTensor(B)], [[0, -1], [-1, 1]],
                       [[0, 1], [0, 1]], add=True)
  return Workload('gemm', [d], fp_type, {'A': A, 'B': B, 'C': C}, {'b': b, 'q': q})


def k_patterns(b=56, density=0.4, seed=1, count=3):
  rng = np.random.default_rng(seed)
  masks = rng.random((count, b, b)) < density
  return [m for m in masks]


def csr_entries(mask):
  """Nonzeros of ``mask`` in row-major (CSR) order."""
  rows, cols = np.nonzero(mask)
  return [(int(r), int(c)) for r, c in zip(rows, cols)]


def ader_volume(fp_type=FloatingPointType.DOUBLE, b=56, q=9, sparse_k=True, density=0.4,
                pattern_seed=1, star=True) -> Workload:
  Q = Tensor([b, q], Addressing.STRIDED, alias='Q')
  Qp = Tensor([b, q], Addressing.STRIDED, alias='Qp')
  tensors = {'Q': Q, 'Qp': Qp}
  masks = k_patterns(b, density, pattern_seed) if sparse_k else [None] * 3
  descrs = []
  for d in range(3):
    if star:
      S = Tensor([q, q], Addressing.STRIDED, spp=ListSPP(star_entries(), (q, q)), alias=f'S{d + 1}')
    else:
      S = Tensor([q, q], Addressing.STRIDED, alias=f'S{d + 1}')
    if masks[d] is not None:
      K = Tensor([b, b], Addressing.NONE, spp=ListSPP(csr_entries(masks[d]), (b, b)),
                 alias=f'K{d + 1}')
    else:
      K = Tensor([b, b], Addressing.NONE, alias=f'K{d + 1}')
    tensors[f'S{d + 1}'] = S
    tensors[f'K{d + 1}'] = K
    qs, ss, ks = SubTensor(Q), SubTensor(S), SubTensor(K)
    T = generate_tmp_tensor([qs, ss], [[0, -1], [-1, 1]])
    descrs.append(MultilinearDescr(SubTensor(T), [qs, ss], [[0, -1], [-1, 1]], [[0, 1], [0, 1]]))
    descrs.append(MultilinearDescr(SubTensor(Qp), [ks, SubTensor(T)], [[0, -1], [-1, 1]],
                                   [[0, 1], [0, 1]], add=d > 0))
  return Workload('ader_volume', descrs, fp_type, tensors,
                  {'b': b, 'q': q, 'sparse_k': sparse_k, 'masks': masks})


def flux_local(fp_type=FloatingPointType.DOUBLE, ordering='mod') -> Workload:
  """Local flux integral (reference ``benchmarks/fused/programs/s3-local-integral.cf``):
  ``i_surf += r_div_m · (f_mr_t · (d_k · a_plus))`` with 64-row padded storage and bounding
  boxes of 56 x 9, 56 x 21, 21 x 56."""
  BB = BoundingBox
  i_surf = Tensor([64, 9], Addressing.STRIDED, BB([0, 0], [56, 9]), alias='i_surf')
  r_div_m = Tensor([64, 32], Addressing.NONE, BB([0, 0], [56, 21]), alias='r_div_m')
  f_mr_t = Tensor([32, 64], Addressing.NONE, BB([0, 0], [21, 56]), alias='f_mr_t')
  d_k = Tensor([64, 9], Addressing.STRIDED, BB([0, 0], [56, 9]), alias='d_k')
  a_plus = Tensor([9, 9], Addressing.STRIDED, BB([0, 0], [9, 9]), alias='a_plus')
  sub = lambda t: SubTensor(t, t.bbox)
  pair = [[0, -1], [-1, 1]]
  ident = [[0, 1], [0, 1]]
  if ordering == 'mod':       # r_div_m · (f_mr_t · (d_k · a_plus))
    t1 = generate_tmp_tensor([sub(d_k), sub(a_plus)], pair)
    t2 = generate_tmp_tensor([sub(f_mr_t), SubTensor(t1)], pair)
    descrs = [MultilinearDescr(SubTensor(t1), [sub(d_k), sub(a_plus)], pair, ident),
              MultilinearDescr(SubTensor(t2), [sub(f_mr_t), SubTensor(t1)], pair, ident)]
  else:                       # r_div_m · ((f_mr_t · d_k) · a_plus)
    t1 = generate_tmp_tensor([sub(f_mr_t), sub(d_k)], pair)
    t2 = generate_tmp_tensor([SubTensor(t1), sub(a_plus)], pair)
    descrs = [MultilinearDescr(SubTensor(t1), [sub(f_mr_t), sub(d_k)], pair, ident),
              MultilinearDescr(SubTensor(t2), [SubTensor(t1), sub(a_plus)], pair, ident)]
  descrs.append(MultilinearDescr(sub(i_surf), [sub(r_div_m), SubTensor(t2)], pair, ident,
                                 add=True))
  tensors = {'i_surf': i_surf, 'r_div_m': r_div_m, 'f_mr_t': f_mr_t, 'd_k': d_k,
             'a_plus': a_plus}
  return Workload('flux_local', descrs, fp_type, tensors, {'ordering': ordering, 'b': 56, 'q': 9})


def flux_neigh(fp_type=FloatingPointType.DOUBLE, ordering='mod') -> Workload:
  """Neighbour flux integral (reference ``benchmarks/fused/programs/s3-neigh-integral.cf``):
  ``dofs += r_div_mat · (f_plus · (r_mat · (i_dofs · a_plus)))``."""
  BB = BoundingBox
  dofs = Tensor([64, 9], Addressing.STRIDED, BB([0, 0], [56, 9]), alias='dofs')
  r_div_mat = Tensor([64, 21], Addressing.NONE, BB([0, 0], [56, 21]), alias='r_div_mat')
  f_plus = Tensor([32, 21], Addressing.NONE, BB([0, 0], [21, 21]), alias='f_plus')
  r_mat = Tensor([32, 56], Addressing.NONE, BB([0, 0], [21, 56]), alias='r_mat')
  i_dofs = Tensor([64, 9], Addressing.STRIDED, BB([0, 0], [56, 9]), alias='i_dofs')
  a_plus = Tensor([9, 9], Addressing.STRIDED, BB([0, 0], [9, 9]), alias='a_plus')
  sub = lambda t: SubTensor(t, t.bbox)
  descrs = []
  if ordering == 'mod':
    t1 = generate_tmp_tensor([sub(i_dofs), sub(a_plus)], [[0, -1], [-1, 1]])
    descrs.append(MultilinearDescr(SubTensor(t1), [sub(i_dofs), sub(a_plus)], [[0, -1], [-1, 1]],
                                   [[0, 1], [0, 1]]))
    t2 = generate_tmp_tensor([sub(r_mat), SubTensor(t1)], [[0, -1], [-1, 1]])
    descrs.append(MultilinearDescr(SubTensor(t2), [sub(r_mat), SubTensor(t1)], [[0, -1], [-1, 1]],
                                   [[0, 1], [0, 1]]))
  else:
    t0 = generate_tmp_tensor([sub(r_mat), sub(i_dofs)], [[0, -1], [-1, 1]])
    descrs.append(MultilinearDescr(SubTensor(t0), [sub(r_mat), sub(i_dofs)], [[0, -1], [-1, 1]],
                                   [[0, 1], [0, 1]]))
    t2 = t0
  t3 = generate_tmp_tensor([sub(f_plus), SubTensor(t2)], [[0, -1], [-1, 1]])
  descrs.append(MultilinearDescr(SubTensor(t3), [sub(f_plus), SubTensor(t2)], [[0, -1], [-1, 1]],
                                 [[0, 1], [0, 1]]))
  if ordering == 'orig':
    t4 = generate_tmp_tensor([SubTensor(t3), sub(a_plus)], [[0, -1], [-1, 1]])
    descrs.append(MultilinearDescr(SubTensor(t4), [SubTensor(t3), sub(a_plus)], [[0, -1], [-1, 1]],
                                   [[0, 1], [0, 1]]))
    t3 = t4
  descrs.append(MultilinearDescr(sub(dofs), [sub(r_div_mat), SubTensor(t3)], [[0, -1], [-1, 1]],
                                 [[0, 1], [0, 1]], add=True))
  tensors = {'dofs': dofs, 'r_div_mat': r_div_mat, 'f_plus': f_plus, 'r_mat': r_mat,
             'i_dofs': i_dofs, 'a_plus': a_plus}
  return Workload('flux_neigh', descrs, fp_type, tensors, {'ordering': ordering, 'b': 56, 'q': 9})


def ader_derivative(fp_type=FloatingPointType.DOUBLE, aligned=True) -> Workload:
  """One step of the ADER time-derivative recursion (reference
  ``benchmarks/fused/programs/seissol-deriv_align.cf`` and ``seissol-deriv_no_align.cf``):
  ``D += C·(A·B) + F·(A'·E) + H·(A''·G)`` with rectangular shared matrices (48×53, 48×54,
  48×55) contracting over the first 53/54/55 rows of 64-row per-element arrays, 9×9 per-element
  star matrices. Aligned: one streamed array ``A`` read from row 0, all 48 rows of ``D``.
  Not aligned: three streamed arrays read from row 1 (the previous derivative's constant mode
  dropped), 35 of 48 rows of ``D`` and of the shared matrices."""
  BB = BoundingBox
  sub = lambda t: SubTensor(t, t.bbox)
  pair, ident = [[0, -1], [-1, 1]], [[0, 1], [0, 1]]
  rows = 48 if aligned else 35
  D = Tensor([48, 9], Addressing.STRIDED, BB([0, 0], [rows, 9]), alias='D')
  tensors = {'D': D}
  descrs = []
  for d, kcols in enumerate((53, 54, 55)):
    if aligned:
      if d == 0:
        tensors['A'] = Tensor([64, 9], Addressing.STRIDED, BB([0, 0], [64, 9]), alias='A')
      A = tensors['A']
    else:
      A = Tensor([64, 9], Addressing.STRIDED, BB([1, 0], [1 + kcols, 9]), alias=f'A{d + 1}')
      tensors[f'A{d + 1}'] = A
    S = Tensor([9, 9], Addressing.STRIDED, BB([0, 0], [9, 9]), alias=f'S{d + 1}')
    K = Tensor([48, kcols], Addressing.NONE, BB([0, 0], [rows, kcols]), alias=f'K{d + 1}')
    tensors[f'S{d + 1}'] = S
    tensors[f'K{d + 1}'] = K
    T = generate_tmp_tensor([sub(A), sub(S)], pair)
    descrs.append(MultilinearDescr(SubTensor(T), [sub(A), sub(S)], pair, ident))
    descrs.append(MultilinearDescr(sub(D), [sub(K), SubTensor(T)], pair, ident, add=True))
  return Workload('ader_derivative', descrs, fp_type, tensors,
                  {'aligned': aligned, 'rows': rows, 'q': 9})


# uneven batching: orders 2..7 (basis sizes) and their share of the elements
MIXED_BASES = (4, 10, 20, 35, 56, 84)
MIXED_WEIGHTS = (0.05, 0.1, 0.2, 0.3, 0.25, 0.1)


def mixed_counts(total=500000, seed=3):
  """Elements per order: ``total`` draws of the order with ``MIXED_WEIGHTS``."""
  rng = np.random.default_rng(seed)
  draw = rng.choice(len(MIXED_BASES), size=total, p=MIXED_WEIGHTS)
  return [int(c) for c in np.bincount(draw, minlength=len(MIXED_BASES))]


def ader_mixed(fp_type=FloatingPointType.DOUBLE, bases=MIXED_BASES):
  """One volume chain per order, dense K_d (b x b), the elastic star pattern."""
  return [ader_volume(fp_type, b=b, sparse_k=False) for b in bases]


# ---------------------------------------------------------------- inputs

def tensor_storage_size(t: Tensor):
  """Entries one element occupies in its array: the dense volume (also for packed sparse
  tensors, whose values sit at the start of each element's slot)."""
  return t.get_real_volume()


def packed_count(t: Tensor):
  return t.spp.count_nz() if t.is_sparse() else t.get_real_volume()


def gemm_inputs(w: Workload, num_elements, seed=0, dtype=np.float64):
  rng = np.random.default_rng(seed)
  b, q = w.meta['b'], w.meta['q']
  A = rng.uniform(-1, 1, size=b * b)
  B = rng.uniform(-1, 1, size=num_elements * b * q)
  C = rng.uniform(-1, 1, size=num_elements * b * q)
  return {'A': A.astype(dtype), 'B': B.astype(dtype), 'C': C.astype(dtype)}


def log_uniform_q(rng, size, lo=-6.0, hi=3.0):
  """Degrees of freedom with magnitudes log-uniform in [10^lo, 10^hi] and random sign."""
  return np.where(rng.random(size) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(lo, hi, size)


def ader_inputs(w: Workload, num_elements, seed=2, dtype=np.float64, q_scale=None):
  """Q ~ N(0,1) (or log-uniform magnitudes when ``q_scale=(lo_exp, hi_exp)``), star values
  U[-2,2] packed at the start of each 81-entry slot, K values U[-1,1] packed in CSR order."""
  rng = np.random.default_rng(seed)
  b, q = w.meta['b'], w.meta['q']
  out = {}
  for d in range(3):
    K = w.tensors[f'K{d + 1}']
    out[f'K{d + 1}'] = rng.uniform(-1, 1, size=packed_count(K)).astype(dtype)
  if q_scale is None:
    out['Q'] = rng.standard_normal(num_elements * b * q).astype(dtype)
  else:
    lo, hi = q_scale
    mag = 10.0 ** rng.uniform(lo, hi, size=num_elements * b * q)
    sign = np.where(rng.random(num_elements * b * q) < 0.5, -1.0, 1.0)
    out['Q'] = (mag * sign).astype(dtype)
  for d in range(3):
    S = w.tensors[f'S{d + 1}']
    nnz = packed_count(S)
    vals = np.zeros((num_elements, tensor_storage_size(S)), dtype=dtype)
    vals[:, :nnz] = rng.uniform(-2, 2, size=(num_elements, nnz))
    out[f'S{d + 1}'] = vals.reshape(-1)
  out['Qp'] = np.zeros(num_elements * b * q, dtype=dtype)
  return out


def launcher_args_by_alias(generator):
  """Launcher operands in order as aliases (``None`` for non-tensor arguments)."""
  return [(name, next((t.alias for t in generator._global_tensors if t.name == name.split('_')[0]), None))
          for _, name in generator.launcher_signature()]


def pointwise_update(fp_type=FloatingPointType.DOUBLE, b=56, q=9) -> Workload:
  Q = Tensor([b, q], Addressing.STRIDED, alias='Q')
  I = Tensor([b, q], Addressing.STRIDED, alias='I')
  dt = Tensor([], Addressing.SCALAR, alias='dt')
  q_var = ot.tensor(Q, indices=[-1, -2])
  stmt = ot.assign(q_var, q_var + ot.scalar(dt) * ot.tensor(I, indices=[-1, -2]))
  return Workload('pointwise_update', [ElementwiseDescr([stmt])], fp_type,
                  {'Q': Q, 'I': I, 'dt': dt}, {'b': b, 'q': q})


def pointwise_plastic(fp_type=FloatingPointType.DOUBLE, b=56, q=9) -> Workload:
  Q = Tensor([b, q], Addressing.STRIDED, alias='Q')
  cohesion = Tensor([], Addressing.SCALAR, alias='cohesion')
  tanphi = Tensor([], Addressing.SCALAR, alias='tanphi')
  s = [ot.tensor(SubTensor(Q, BoundingBox([0, k], [b, k + 1])), indices=[-1, -2])
       for k in range(6)]
  m, tau, yld, f = ot.temp(), ot.temp(), ot.temp(), ot.temp()
  stmts = [
    ot.assign(m, (s[0] + s[1] + s[2]) / 3.0),
    ot.assign(tau, ot.sqrt(0.5 * ((s[0] - m) ** 2 + (s[1] - m) ** 2 + (s[2] - m) ** 2)
                           + s[3] ** 2 + s[4] ** 2 + s[5] ** 2)),
    ot.assign(yld, ot.scalar(cohesion) - m * ot.scalar(tanphi)),
    ot.conditional(tau > yld, [ot.assign(f, yld / tau)] +
                   [ot.assign(s[k], m + f * (s[k] - m)) for k in range(3)] +
                   [ot.assign(s[k], f * s[k]) for k in range(3, 6)]),
  ]
  return Workload('pointwise_plastic', [ElementwiseDescr(stmts)], fp_type,
                  {'Q': Q, 'cohesion': cohesion, 'tanphi': tanphi}, {'b': b, 'q': q})
