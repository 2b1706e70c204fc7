"""Pins the FP64 oracle (oracle/chain_oracle.py) before anything is compared against it.

Each check uses a formulation independent of the oracle's generic einsum interpreter: batched
``matmul`` closed forms, hand-computed tiny cases, brute-force loops, and invariants chosen so a
dropped term, a sign slip, a transposed operand or a wrong offset fails at least one of them.
"""
import numpy as np
import pytest

from oracle.chain_oracle import OracleInputs, run
from tensorforge_amd import workloads as W
from tensorforge_amd.common.basic_types import Addressing, FloatingPointType
from tensorforge_amd.common.matrix.boundingbox import BoundingBox
from tensorforge_amd.common.matrix.spp import ListSPP
from tensorforge_amd.common.matrix.tensor import SubTensor, Tensor
from tensorforge_amd.generators.descriptions import GemmDescr, MultilinearDescr
from tensorforge_amd.generators.generator import Generator
from tensorforge_amd.common.context import Context


def cm(x, e, r, c):
  """Column-major per-element view: flat (e*r*c) -> (e, r, c)."""
  return x.reshape(e, c, r).transpose(0, 2, 1)


def _named(descrs):
  Generator(descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE)).register()


def test_gemm_closed_form():
  w = W.gemm()
  _named(w.descrs)
  e = 6
  x = W.gemm_inputs(w, e, seed=0)
  t = w.tensors
  out = run(w.descrs, OracleInputs({t['A']: x['A'], t['B']: x['B'], t['C']: x['C'].copy()}, e))
  A = x['A'].reshape(56, 56, order='F')
  ref = cm(x['C'], e, 56, 9) + A @ cm(x['B'], e, 56, 9)
  np.testing.assert_allclose(cm(out[t['C']], e, 56, 9), ref, rtol=1e-14, atol=1e-13)


def test_tiny_hand_computed():
  # C(2x1) += A(2x2) B(2x1): [[1,2],[3,4]] [5,6]^T + [1,1] = [18, 40]
  A = Tensor([2, 2], Addressing.NONE)
  B = Tensor([2, 1], Addressing.STRIDED)
  C = Tensor([2, 1], Addressing.STRIDED)
  d = MultilinearDescr(SubTensor(C), [SubTensor(A), SubTensor(B)], [[0, -1], [-1, 1]],
                       [[0, 1], [0, 1]], add=True)
  _named([d])
  out = run([d], OracleInputs({A: np.array([1., 3., 2., 4.]), B: np.array([5., 6.]),
                               C: np.array([1., 1.])}, 1))
  np.testing.assert_array_equal(out[C], [18.0, 40.0])


def test_ader_closed_form_sparse_and_dense():
  for sparse in (True, False):
    w = W.ader_volume(sparse_k=sparse)
    _named(w.descrs)
    e = 4
    x = W.ader_inputs(w, e)
    t = w.tensors
    out = run(w.descrs, OracleInputs({t[k]: v.copy() for k, v in x.items()}, e))
    Q = cm(x['Q'], e, 56, 9)
    ref = np.zeros_like(Q)
    for d in range(3):
      K = np.zeros((56, 56))
      if sparse:
        for pos, (r, c) in enumerate(W.csr_entries(w.meta['masks'][d])):
          K[r, c] = x[f'K{d + 1}'][pos]
      else:
        K = x[f'K{d + 1}'].reshape(56, 56, order='F')
      S = np.zeros((e, 9, 9))
      sv = x[f'S{d + 1}'].reshape(e, 81)
      for pos, (r, c) in enumerate(W.star_entries()):
        S[:, r, c] = sv[:, pos]
      ref += K @ Q @ S
    np.testing.assert_allclose(cm(out[t['Qp']], e, 56, 9), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('ta,tb', [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_descr_transposes_alpha_beta(ta, tb):
  rng = np.random.default_rng(3)
  m, n, k, e = 5, 4, 3, 3
  A = Tensor([k, m] if ta else [m, k], Addressing.NONE)
  B = Tensor([n, k] if tb else [k, n], Addressing.STRIDED)
  C = Tensor([m, n], Addressing.STRIDED)
  d = GemmDescr(ta, tb, SubTensor(A), SubTensor(B), SubTensor(C), alpha=1.5, beta=0.5)
  _named([d])
  a = rng.standard_normal(m * k)
  b = rng.standard_normal(e * n * k)
  c = rng.standard_normal(e * m * n)
  out = run([d], OracleInputs({A: a, B: b, C: c.copy()}, e))
  Am = a.reshape((k, m) if ta else (m, k), order='F')
  Bm = cm(b, e, n, k) if tb else cm(b, e, k, n)
  opA = Am.T if ta else Am
  opB = Bm.transpose(0, 2, 1) if tb else Bm
  ref = 1.5 * opA @ opB + 0.5 * cm(c, e, m, n)
  np.testing.assert_allclose(cm(out[C], e, m, n), ref, rtol=1e-13, atol=1e-13)


def test_three_index_brute_force():
  rng = np.random.default_rng(4)
  A = Tensor([3, 2], Addressing.NONE)
  B = Tensor([2, 2, 3], Addressing.STRIDED)
  C = Tensor([3, 2, 3], Addressing.STRIDED)
  d = MultilinearDescr(SubTensor(C), [SubTensor(A), SubTensor(B)], [[0, -1], [-1, 1, 2]],
                       [[0, 1], [0, 1, 2]])
  _named([d])
  a, b = rng.standard_normal(6), rng.standard_normal(2 * 12)
  out = run([d], OracleInputs({A: a, B: b, C: np.zeros(2 * 18)}, 2))
  for el in range(2):
    for i in range(3):
      for j in range(2):
        for kk in range(3):
          s = sum(a[i + 3 * l] * b[el * 12 + l + 2 * j + 4 * kk] for l in range(2))
          assert abs(out[C][el * 18 + i + 3 * j + 6 * kk] - s) < 1e-14


def test_flags_offsets_and_pointer_based_invariants():
  rng = np.random.default_rng(5)
  e = 7
  A = Tensor([4, 3], Addressing.NONE)
  B = Tensor([3, 2], Addressing.STRIDED)
  C = Tensor([4, 2], Addressing.STRIDED)
  d = MultilinearDescr(SubTensor(C), [SubTensor(A), SubTensor(B)], [[0, -1], [-1, 1]],
                       [[0, 1], [0, 1]], add=True)
  _named([d])
  a = rng.standard_normal(12)
  b = rng.standard_normal(e * 6 + 5)
  c = rng.standard_normal(e * 8 + 3)
  flags = np.array([1, 0, 1, 1, 0, 1, 1], dtype=np.uint32)
  out = run([d], OracleInputs({A: a, B: b, C: c.copy()}, e, flags=flags, offsets={B: 5, C: 3}))
  Am = a.reshape(4, 3, order='F')
  for el in range(e):
    cb = c[3 + el * 8: 3 + el * 8 + 8].reshape(4, 2, order='F')
    bb = b[5 + el * 6: 5 + el * 6 + 6].reshape(3, 2, order='F')
    want = cb + Am @ bb if flags[el] else cb
    np.testing.assert_allclose(out[C][3 + el * 8: 3 + el * 8 + 8].reshape(4, 2, order='F'), want,
                               rtol=1e-14, atol=1e-14)
  np.testing.assert_array_equal(out[C][:3], c[:3])
  # pointer-based operands give the same result as strided ones
  Bp = Tensor([3, 2], Addressing.PTR_BASED)
  Cp = Tensor([4, 2], Addressing.PTR_BASED)
  d2 = MultilinearDescr(SubTensor(Cp), [SubTensor(A), SubTensor(Bp)], [[0, -1], [-1, 1]],
                        [[0, 1], [0, 1]], add=True)
  _named([d2])
  bl = [b[5 + el * 6: 5 + el * 6 + 6].copy() for el in range(e)]
  cl = [c[3 + el * 8: 3 + el * 8 + 8].copy() for el in range(e)]
  out2 = run([d2], OracleInputs({A: a, Bp: bl, Cp: cl}, e, flags=flags))
  for el in range(e):
    np.testing.assert_allclose(out2[Cp][el], out[C][3 + el * 8: 3 + el * 8 + 8], rtol=0, atol=0)


def test_linearity_and_sparse_equals_densified():
  rng = np.random.default_rng(6)
  mask = rng.random((6, 5)) < 0.4
  entries = W.csr_entries(mask)
  As = Tensor([6, 5], Addressing.NONE, spp=ListSPP(entries, (6, 5)))
  Ad = Tensor([6, 5], Addressing.NONE)
  B = Tensor([5, 3], Addressing.STRIDED)
  C = Tensor([6, 3], Addressing.STRIDED)
  ds = MultilinearDescr(SubTensor(C), [SubTensor(As), SubTensor(B)], [[0, -1], [-1, 1]],
                        [[0, 1], [0, 1]])
  _named([ds])
  vals = rng.standard_normal(len(entries))
  dense = np.zeros((6, 5))
  for p, (r, c) in enumerate(entries):
    dense[r, c] = vals[p]
  b1, b2 = rng.standard_normal(2 * 15), rng.standard_normal(2 * 15)
  o1 = run([ds], OracleInputs({As: vals, B: b1, C: np.zeros(36)}, 2))[C].copy()
  o2 = run([ds], OracleInputs({As: vals, B: b2, C: np.zeros(36)}, 2))[C].copy()
  o3 = run([ds], OracleInputs({As: vals, B: 2 * b1 - b2, C: np.zeros(36)}, 2))[C].copy()
  np.testing.assert_allclose(o3, 2 * o1 - o2, rtol=1e-13, atol=1e-13)
  dd = MultilinearDescr(SubTensor(C), [SubTensor(Ad), SubTensor(B)], [[0, -1], [-1, 1]],
                        [[0, 1], [0, 1]])
  _named([dd])
  od = run([dd], OracleInputs({Ad: dense.reshape(-1, order='F'), B: b1, C: np.zeros(36)}, 2))[C]
  np.testing.assert_allclose(o1, od, rtol=1e-14, atol=1e-14)


def test_bbox_window_reads_only_its_window():
  rng = np.random.default_rng(9)
  A = Tensor([8, 6], Addressing.NONE, BoundingBox([2, 1], [7, 5]))
  B = Tensor([6, 3], Addressing.STRIDED, BoundingBox([1, 0], [5, 3]))
  C = Tensor([8, 3], Addressing.STRIDED, BoundingBox([2, 0], [7, 3]))
  d = MultilinearDescr(SubTensor(C), [SubTensor(A), SubTensor(B)], [[0, -1], [-1, 1]],
                       [[0, 1], [0, 1]], add=True)
  _named([d])
  a, b, c = rng.standard_normal(48), rng.standard_normal(18), rng.standard_normal(24)
  out = run([d], OracleInputs({A: a, B: b, C: c.copy()}, 1))
  # A's stored origin is (2,1): entry (i,k) at (i-2) + (k-1)*8
  for i in range(2, 7):
    for j in range(3):
      s = sum(a[(i - 2) + (k - 1) * 8] * b[(k - 1) + j * 6] for k in range(1, 5))
      assert abs(out[C][(i - 2) + j * 8] - (c[(i - 2) + j * 8] + s)) < 1e-13


def test_omp_restatement_matches_oracle():
  import ctypes
  import os
  lib_path = os.path.join(os.path.dirname(os.path.dirname(__file__)), 'oracle', 'lib',
                          'libomp_chain.so')
  if not os.path.exists(lib_path):
    pytest.skip('oracle C library not built')
  lib = ctypes.CDLL(lib_path)
  w = W.gemm()
  _named(w.descrs)
  e = 16
  x = W.gemm_inputs(w, e)
  t = w.tensors
  ref = run(w.descrs, OracleInputs({t['A']: x['A'], t['B']: x['B'], t['C']: x['C'].copy()}, e))[t['C']]
  P = ctypes.POINTER(ctypes.c_double)
  for fn, extra in ((lib.omp_gemm_flat, ()), (lib.omp_gemm_team, (32, 8))):
    c = x['C'].copy()
    fn(x['A'].ctypes.data_as(P), x['B'].ctypes.data_as(P), c.ctypes.data_as(P), 56, 9,
       ctypes.c_long(e), *extra)
    np.testing.assert_allclose(c, ref, rtol=1e-13, atol=1e-13)
