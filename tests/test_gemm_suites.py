"""The oracle on the reference's 514 dense GEMM test cases (tests/gemm_suites.py). CPU only.

The suites were written for an older interface in which every bounding box was relative to its
own corner (reference tests/gemms/generate.py:83-140: ``C_sub = alpha op(A_sub) op(B_sub) +
beta C_sub``). The current ``MultilinearDescr`` semantics ranges each index over the intersection
of the bounding boxes it indexes, in absolute coordinates
(backend/instructions/compute/multilinear.py:58-96). Where each index's boxes coincide (26 cases)
both agree and the oracle is checked against the reference driver's closed form; elsewhere
(488 cases: partial and empty overlaps) against the intersection written with NumPy slices.
"""
import numpy as np
import pytest

import gemm_suites as gs
from oracle.chain_oracle import OracleInputs, run as oracle_run


def driver_closed_form(case, host):
  """The reference driver's check, element by element, on full-matrix arrays."""
  out = host['C'].copy()
  (ra, ca), (rb, cb), (rc, cc) = case['a'], case['b'], case['c']
  for e in range(case['n']):
    A = host['A'][e * ra * ca:(e + 1) * ra * ca].reshape(ca, ra).T
    B = host['B'][e * rb * cb:(e + 1) * rb * cb].reshape(cb, rb).T
    C = out[e * rc * cc:(e + 1) * rc * cc].reshape(cc, rc).T     # a view into out
    a0, a1, a2, a3 = case['bbox_a']
    b0, b1, b2, b3 = case['bbox_b']
    c0, c1, c2, c3 = case['bbox_c']
    opA = A[a0:a2, a1:a3].T if case['trans_a'] else A[a0:a2, a1:a3]
    opB = B[b0:b2, b1:b3].T if case['trans_b'] else B[b0:b2, b1:b3]
    C[c0:c2, c1:c3] = case['alpha'] * (opA @ opB) + case['beta'] * C[c0:c2, c1:c3]
  return out


def index_ranges(case):
  """Absolute ranges of the output rows (m), the contraction (k) and the output columns (n) as
  each operand sees them."""
  a0, a1, a2, a3 = case['bbox_a']
  b0, b1, b2, b3 = case['bbox_b']
  c0, c1, c2, c3 = case['bbox_c']
  m_a, k_a = ((a1, a3), (a0, a2)) if case['trans_a'] else ((a0, a2), (a1, a3))
  k_b, n_b = ((b1, b3), (b0, b2)) if case['trans_b'] else ((b0, b2), (b1, b3))
  return dict(m=[m_a, (c0, c2)], k=[k_a, k_b], n=[n_b, (c1, c3)])


def aligned(case):
  return all(len(set(v)) == 1 for v in index_ranges(case).values())


def intersection_closed_form(case, host):
  """C[i, j] = alpha sum_k op(A)[i, k] op(B)[k, j] + beta C[i, j] for i, j, k in the
  intersections, absolute coordinates; entries outside are untouched."""
  out = host['C'].copy()
  (ra, ca), (rb, cb), (rc, cc) = case['a'], case['b'], case['c']
  r = {key: (max(lo for lo, _ in v), min(hi for _, hi in v)) for key, v in index_ranges(case).items()}
  (m0, m1), (k0, k1), (n0, n1) = r['m'], r['k'], r['n']
  if m1 <= m0 or n1 <= n0:
    return out
  for e in range(case['n']):
    A = host['A'][e * ra * ca:(e + 1) * ra * ca].reshape(ca, ra).T
    B = host['B'][e * rb * cb:(e + 1) * rb * cb].reshape(cb, rb).T
    C = out[e * rc * cc:(e + 1) * rc * cc].reshape(cc, rc).T
    opA = A.T if case['trans_a'] else A
    opB = B.T if case['trans_b'] else B
    prod = opA[m0:m1, k0:k1] @ opB[k0:k1, n0:n1] if k1 > k0 else 0.0
    C[m0:m1, n0:n1] = case['alpha'] * prod + case['beta'] * C[m0:m1, n0:n1]
  return out


def host_inputs(case, seed):
  rng = np.random.default_rng(seed)
  return {k: rng.uniform(-1, 1, size=case['n'] * case[k.lower()][0] * case[k.lower()][1])
          for k in ('A', 'B', 'C')}


def offsets(case):
  return {k: gs.offset(case[k.lower()], case[f'bbox_{k.lower()}']) for k in ('A', 'B', 'C')}


def expected(case, host):
  return driver_closed_form(case, host) if aligned(case) else intersection_closed_form(case, host)


def test_case_count_and_dimensions():
  cases = gs.cases()
  assert len(cases) == 514
  assert sum(aligned(c) for c in cases) == 26
  for c in cases:
    def dims(bbox):
      return bbox[2] - bbox[0], bbox[3] - bbox[1]
    ar, ac = dims(c['bbox_a'])
    br, bc = dims(c['bbox_b'])
    cr, cc = dims(c['bbox_c'])
    m, k = (ac, ar) if c['trans_a'] else (ar, ac)
    k2, n = (bc, br) if c['trans_b'] else (br, bc)
    assert (m, n) == (cr, cc) and k == k2, c['name']


@pytest.mark.parametrize('index', range(0, 514, 1))
def test_oracle_matches_reference_driver(index):
  case = gs.cases()[index]
  descrs, tensors = gs.descriptions(case)
  host = host_inputs(case, index)
  want = expected(case, host)
  offs = offsets(case)
  inp = OracleInputs({tensors[k]: host[k].copy() for k in host}, case['n'],
                     offsets={tensors[k]: offs[k] for k in offs})
  got = oracle_run(descrs, inp)[tensors['C']]
  np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


# ---- sparse-operand and scale-add suites ----
def sparse_host(case, seed):
  rng = np.random.default_rng(seed)
  m = case['a'][1] if case['trans_a'] else case['a'][0]
  n = case['b'][0] if case['trans_b'] else case['b'][1]
  host = {k: rng.uniform(-1, 1, size=case['n'] * r * c)
          for k, (r, c) in (('A', case['a']), ('B', case['b']), ('C', (m, n)))}
  return host, (m, n)


def sparse_closed_form(case, host, mn):
  """op(A) op(B) + C per element, the sparse operand densified from its packed values (the
  pattern's list order, column-major as the reference's loader sorts it)."""
  out = host['C'].copy()
  m, n = mn
  vol = {k: case[k][0] * case[k][1] for k in ('a', 'b')}
  for e in range(case['n']):
    mats = {}
    for k in ('a', 'b'):
      r, c = case[k]
      flat = host[k.upper()][e * vol[k]:(e + 1) * vol[k]]
      if case['sparse'] == k:
        dense = np.zeros((r, c))
        for pos, (i, j) in enumerate(case['pattern']):
          dense[i, j] = flat[pos]
      else:
        dense = flat.reshape(c, r).T
      mats[k] = dense
    opA = mats['a'].T if case['trans_a'] else mats['a']
    opB = mats['b'].T if case['trans_b'] else mats['b']
    C = out[e * m * n:(e + 1) * m * n].reshape(n, m).T
    C[:, :] = opA @ opB + C
  return out


@pytest.mark.parametrize('index', range(len(gs.SPARSE_SUITES)))
def test_sparse_suites_oracle(index):
  case = gs.sparse_cases()[index]
  descrs, tensors = gs.sparse_descriptions(case)
  host, mn = sparse_host(case, index)
  want = sparse_closed_form(case, host, mn)
  inp = OracleInputs({tensors[k]: host[k].copy() for k in host}, case['n'])
  got = oracle_run(descrs, inp)[tensors['C']]
  np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def csa_host(case, seed):
  rng = np.random.default_rng(seed)
  return {k: rng.uniform(-1, 1, size=case['n'] * case[k.lower()][0] * case[k.lower()][1])
          for k in ('A', 'C')}


def csa_closed_form(case, host):
  out = host['C'].copy()
  (ra, ca), (rc, cc) = case['a'], case['c']
  r0, c0, r1, c1 = case['bbox']
  for e in range(case['n']):
    A = host['A'][e * ra * ca:(e + 1) * ra * ca].reshape(ca, ra).T
    C = out[e * rc * cc:(e + 1) * rc * cc].reshape(cc, rc).T
    C[r0:r1, c0:c1] = case['alpha'] * A[r0:r1, c0:c1] + case['beta'] * C[r0:r1, c0:c1]
  return out


@pytest.mark.parametrize('index', range(5))
def test_csa_suites_oracle(index):
  from oracle.pointwise_oracle import run as pw_run
  case = gs.csa_cases()[index]
  descrs, tensors = gs.csa_descriptions(case)
  host = csa_host(case, index)
  want = csa_closed_form(case, host)
  got = pw_run(descrs, OracleInputs({tensors[k]: host[k].copy() for k in host}, case['n']))
  np.testing.assert_allclose(got[tensors['C']], want, rtol=0, atol=1e-15)
