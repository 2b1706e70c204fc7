"""Description lists whose kernel names and launcher prototypes are pinned by golden vectors.

Written against the ``tensorforge`` description API only (``Tensor``, ``SubTensor``,
``BoundingBox``, ``ListSPP``, ``MultilinearDescr``, ``GemmDescr``, ``generate_tmp_tensor``), so
the same cases can be built with any package that provides that API: ``make_kernel_names.py``
builds them with the reference generator to write ``kernel_names.json``; the tests build them
with this package and compare.

Each case is ``(label, fp_type_name, descr_list)``; tensors are created inside the case so that
nothing is shared between cases.
"""
import numpy as np


def _csr(mask):
  rows, cols = np.nonzero(mask)
  return [(int(r), int(c)) for r, c in zip(rows, cols)]


STAR = sorted([(r, c) for r in range(3) for c in range(6, 9)] +
              [(3, 6), (3, 7), (4, 7), (4, 8), (5, 6), (5, 8)] +
              [(6, 0), (6, 3), (6, 5), (7, 3), (7, 1), (7, 4), (8, 5), (8, 4), (8, 2)],
              key=lambda rc: (rc[1], rc[0]))


def cases():
  from tensorforge.common.aux import generate_tmp_tensor
  from tensorforge.common.basic_types import Addressing
  from tensorforge.common.matrix.boundingbox import BoundingBox
  from tensorforge.common.matrix.spp import ListSPP
  from tensorforge.common.matrix.tensor import SubTensor, Tensor
  from tensorforge.generators.descriptions import GemmDescr, MultilinearDescr

  pair = [[0, -1], [-1, 1]]
  ident = [[0, 1], [0, 1]]
  out = []

  # single batched GEMM, C_e += A B_e (A shared)
  A = Tensor([56, 56], Addressing.NONE, alias='A')
  B = Tensor([56, 9], Addressing.STRIDED, alias='B')
  C = Tensor([56, 9], Addressing.STRIDED, alias='C')
  out.append(('gemm_56x9_add', 'double',
              [MultilinearDescr(SubTensor(C), [SubTensor(A), SubTensor(B)], pair, ident,
                                add=True)]))

  # ADER volume chain: Q' = sum_d K_d Q S_d, sparse K (CSR order) and star-pattern S
  def ader(sparse_k, b=56, q=9):
    rng = np.random.default_rng(1)
    masks = rng.random((3, b, b)) < 0.4
    Q = Tensor([b, q], Addressing.STRIDED, alias='Q')
    Qp = Tensor([b, q], Addressing.STRIDED, alias='Qp')
    descrs = []
    for d in range(3):
      S = Tensor([q, q], Addressing.STRIDED, spp=ListSPP(STAR, (q, q)), alias=f'S{d + 1}')
      if sparse_k:
        K = Tensor([b, b], Addressing.NONE, spp=ListSPP(_csr(masks[d]), (b, b)),
                   alias=f'K{d + 1}')
      else:
        K = Tensor([b, b], Addressing.NONE, alias=f'K{d + 1}')
      qs, ss, ks = SubTensor(Q), SubTensor(S), SubTensor(K)
      T = generate_tmp_tensor([qs, ss], pair)
      descrs.append(MultilinearDescr(SubTensor(T), [qs, ss], pair, ident))
      descrs.append(MultilinearDescr(SubTensor(Qp), [ks, SubTensor(T)], pair, ident,
                                     add=d > 0))
    return descrs
  out.append(('ader_volume_o6_sparse', 'double', ader(True)))
  out.append(('ader_volume_o6_dense', 'double', ader(False)))
  out.append(('ader_volume_o4_dense_float', 'float', ader(False, b=20)))

  # local flux integral with padded storage and bounding boxes
  BB = BoundingBox
  i_surf = Tensor([64, 9], Addressing.STRIDED, BB([0, 0], [56, 9]), alias='i_surf')
  r_div_m = Tensor([64, 32], Addressing.NONE, BB([0, 0], [56, 21]), alias='r_div_m')
  f_mr_t = Tensor([32, 64], Addressing.NONE, BB([0, 0], [21, 56]), alias='f_mr_t')
  d_k = Tensor([64, 9], Addressing.STRIDED, BB([0, 0], [56, 9]), alias='d_k')
  a_plus = Tensor([9, 9], Addressing.STRIDED, BB([0, 0], [9, 9]), alias='a_plus')
  sub = lambda t: SubTensor(t, t.bbox)
  t1 = generate_tmp_tensor([sub(d_k), sub(a_plus)], pair)
  t2 = generate_tmp_tensor([sub(f_mr_t), SubTensor(t1)], pair)
  out.append(('flux_local_bbox', 'double',
              [MultilinearDescr(SubTensor(t1), [sub(d_k), sub(a_plus)], pair, ident),
               MultilinearDescr(SubTensor(t2), [sub(f_mr_t), SubTensor(t1)], pair, ident),
               MultilinearDescr(sub(i_surf), [sub(r_div_m), SubTensor(t2)], pair, ident,
                                add=True)]))

  # pointer-based operands and a sub-block window
  P = Tensor([24, 20], Addressing.PTR_BASED, alias='P')
  R = Tensor([20, 9], Addressing.STRIDED, alias='R')
  D = Tensor([32, 9], Addressing.PTR_BASED, BB([4, 0], [28, 9]), alias='D')
  out.append(('ptr_based_window', 'double',
              [MultilinearDescr(SubTensor(D, BB([4, 0], [28, 9])), [SubTensor(P), SubTensor(R)],
                                pair, ident)]))

  # GemmDescr: transposes (no alpha, no beta: the forms in which both packages agree)
  Ga = Tensor([20, 24], Addressing.STRIDED, alias='Ga')
  Gb = Tensor([9, 20], Addressing.NONE, alias='Gb')
  Gc = Tensor([24, 9], Addressing.STRIDED, alias='Gc')
  out.append(('gemm_descr_trans_ab', 'double',
              [GemmDescr(True, True, SubTensor(Ga), SubTensor(Gb), SubTensor(Gc))]))

  # sparse x dense and dense x sparse (patterns from a seeded mask, packed in pattern order)
  rng = np.random.default_rng(7)
  sp_a = Tensor([16, 12], Addressing.STRIDED, spp=ListSPP(_csr(rng.random((16, 12)) < 0.3),
                                                          (16, 12)), alias='SA')
  dn_b = Tensor([12, 10], Addressing.NONE, alias='DB')
  dn_c = Tensor([16, 10], Addressing.STRIDED, alias='DC')
  out.append(('sparse_dense', 'double',
              [MultilinearDescr(SubTensor(dn_c), [SubTensor(sp_a), SubTensor(dn_b)], pair, ident,
                                add=True)]))
  dn_a = Tensor([16, 12], Addressing.PTR_BASED, alias='DA')
  sp_b = Tensor([12, 10], Addressing.NONE, spp=ListSPP(_csr(rng.random((12, 10)) < 0.3),
                                                       (12, 10)), alias='SB')
  dn_d = Tensor([16, 10], Addressing.PTR_BASED, alias='DD')
  out.append(('dense_sparse', 'float',
              [MultilinearDescr(SubTensor(dn_d), [SubTensor(dn_a), SubTensor(sp_b)], pair,
                                ident)]))

  # a by-value scalar factor as a third operand, and GemmDescr's numeric alpha
  S1 = Tensor([], Addressing.SCALAR, alias='s')
  Ma = Tensor([10, 7], Addressing.STRIDED, alias='Ma')
  Mb = Tensor([7, 5], Addressing.STRIDED, alias='Mb')
  Mc = Tensor([10, 5], Addressing.STRIDED, alias='Mc')
  out.append(('scalar_factor', 'double',
              [MultilinearDescr(SubTensor(Mc), [SubTensor(Ma), SubTensor(Mb), SubTensor(S1)],
                                [[0, -1], [-1, 1], []], [[0, 1], [0, 1], []], add=True)]))
  Na = Tensor([10, 7], Addressing.NONE, alias='Na')
  Nb = Tensor([7, 5], Addressing.STRIDED, alias='Nb')
  Nc = Tensor([10, 5], Addressing.STRIDED, alias='Nc')
  out.append(('gemm_descr_alpha', 'double',
              [GemmDescr(False, False, SubTensor(Na), SubTensor(Nb), SubTensor(Nc), alpha=2.5)]))

  # three-index contraction: D[i,j,l] = A[i,k] B[k,j,l]
  A3 = Tensor([8, 6], Addressing.NONE, alias='A3')
  B3 = Tensor([6, 5, 4], Addressing.STRIDED, alias='B3')
  D3 = Tensor([8, 5, 4], Addressing.STRIDED, alias='D3')
  out.append(('three_index', 'float',
              [MultilinearDescr(SubTensor(D3), [SubTensor(A3), SubTensor(B3)],
                                [[0, -1], [-1, 1, 2]], [[0, 1], [0, 1, 2]])]))
  return out


def aliases(descrs):
  """Aliases of the named tensors of a description list, in first-appearance order."""
  out = []
  for d in descrs:
    for sub in [d.dest] + list(d.ops):
      alias = getattr(getattr(sub, 'tensor', sub), 'alias', None)
      if alias is not None and alias not in out:
        out.append(alias)
  return out


def call_site_maps(descrs):
  """Name maps for ``Generator.generate_call_site``: ``<alias>_ptr`` / ``<alias>_off``."""
  names = aliases(descrs)
  return ({a: f'{a}_ptr' for a in names}, {a: f'{a}_off' for a in names})


def extra_cases():
  """Cases beyond the name set whose numbers are pinned by the OpenMP fixtures
  (``make_omp_fixtures.py``): the ADER volume chain in FP32 and at order 7 (84 basis functions,
  dense K, the precision study's shape)."""
  from tensorforge.common.aux import generate_tmp_tensor
  from tensorforge.common.basic_types import Addressing
  from tensorforge.common.matrix.spp import ListSPP
  from tensorforge.common.matrix.tensor import SubTensor, Tensor
  from tensorforge.generators.descriptions import MultilinearDescr

  pair = [[0, -1], [-1, 1]]
  ident = [[0, 1], [0, 1]]

  def ader(sparse_k, b, q=9):
    rng = np.random.default_rng(1)
    masks = rng.random((3, b, b)) < 0.4
    Q = Tensor([b, q], Addressing.STRIDED, alias='Q')
    Qp = Tensor([b, q], Addressing.STRIDED, alias='Qp')
    descrs = []
    for d in range(3):
      S = Tensor([q, q], Addressing.STRIDED, spp=ListSPP(STAR, (q, q)), alias=f'S{d + 1}')
      if sparse_k:
        K = Tensor([b, b], Addressing.NONE, spp=ListSPP(_csr(masks[d]), (b, b)),
                   alias=f'K{d + 1}')
      else:
        K = Tensor([b, b], Addressing.NONE, alias=f'K{d + 1}')
      qs, ss, ks = SubTensor(Q), SubTensor(S), SubTensor(K)
      T = generate_tmp_tensor([qs, ss], pair)
      descrs.append(MultilinearDescr(SubTensor(T), [qs, ss], pair, ident))
      descrs.append(MultilinearDescr(SubTensor(Qp), [ks, SubTensor(T)], pair, ident,
                                     add=d > 0))
    return descrs

  return [('ader_volume_o6_sparse_float', 'float', ader(True, 56)),
          ('ader_volume_o7_dense', 'double', ader(False, 84))]
