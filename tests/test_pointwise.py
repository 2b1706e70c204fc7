"""Pointwise statements (ElementwiseDescr): the optree API, generation, and the pointwise oracle
pinned by known answers, closed forms and invariants. CPU only."""
import math

import numpy as np
import pytest

from oracle.chain_oracle import OracleInputs
from oracle.pointwise_oracle import run as pw_run
from tensorforge_amd import catalog
from tensorforge_amd.common.basic_types import Addressing, DataFlowDirection, FloatingPointType
from tensorforge_amd.common.context import Context
from tensorforge_amd.common.exceptions import GenerationError
from tensorforge_amd.common.matrix.spp import ListSPP
from tensorforge_amd.common.matrix.tensor import SubTensor, Tensor
from tensorforge_amd.common.operation import Operation
from tensorforge_amd.emit import pointwise as pw
from tensorforge_amd.generators import optree as ot
from tensorforge_amd.generators.descriptions import ElementwiseDescr, MultilinearDescr
from tensorforge_amd.generators.generator import Generator

F64 = FloatingPointType.DOUBLE


def _gen(descrs, fp=F64):
  g = Generator(descrs, Context('gfx950', 'hip', fp))
  g.generate()
  return g


def _x():
  X = Tensor([4, 3], Addressing.STRIDED, alias='X')
  return X, ot.tensor(X, indices=[-1, -2])


# ---- the description API ----
def test_builders_rewrite_trivial_forms():
  _, x = _x()
  assert ot.mul(x, 1) is x and ot.mul(1.0, x) is x
  assert ot.mul(x, -1).optype == Operation.NEG
  assert (x / 1) is x and (x / -1).optype == Operation.NEG
  assert (1 / x).optype == Operation.RCP
  sq = x ** 2
  assert sq.optype == Operation.MUL and sq.operands[0] is x and sq.operands[1] is x
  assert (x ** 0.5).optype == Operation.SQRT
  assert (x ** -0.5).optype == Operation.RSQRT
  assert (x ** (1 / 3)).optype == Operation.CBRT
  assert (x ** (-1 / 3)).optype == Operation.RCBRT
  assert (x ** -1).optype == Operation.RCP
  assert (x ** 1) is x
  assert (math.e ** x).optype == Operation.EXP
  assert (x ** 1.7).optype == Operation.POW
  assert ot.sinh(x).optype == Operation.SINH != Operation.SIN


def test_operators_build_nodes():
  _, x = _x()
  cases = {x + 1: Operation.ADD, 2 - x: Operation.SUB, x * x: Operation.MUL, x / x: Operation.DIV,
           x % 3: Operation.MOD, x < 1: Operation.LT, x <= 1: Operation.LE, x > 1: Operation.GT,
           x >= 1: Operation.GE, x == 1: Operation.EQ, x != 1: Operation.NEQ, -x: Operation.NEG,
           abs(x): Operation.ABS, (x > 1) & (x < 2): Operation.AND,
           (x > 1) | (x < 2): Operation.OR, ~(x > 1): Operation.NOT}
  for node, op in cases.items():
    assert node.optype == op


def test_descr_directions_names_and_temps():
  w = catalog.get('pw_plastic_f64')
  d = w.descrs[0]
  assert w.tensors['Q'].direction == DataFlowDirection.SOURCESINK
  assert w.tensors['cohesion'].direction == DataFlowDirection.SOURCE
  assert 'tmp0' in str(d) and 'tmp3' in str(d)
  # deterministic kernel names: the same description built twice hashes the same
  a = catalog.get('pw_plastic_f64').generator
  b = catalog.get('pw_plastic_f64').generator
  a.register()
  b.register()
  assert a.get_base_name() == b.get_base_name()
  assert str(catalog.get('t_pw_types').descrs[0]) == str(catalog.get('t_pw_types').descrs[0])


def test_launcher_signature_follows_first_appearance():
  g = catalog.get('pw_update_f64').generator
  g.register()
  sig = g.launcher_signature()
  # Q is read first (Q + dt*I), then dt, then I; the destination Q is already named
  assert [n for _, n in sig] == ['m0', 'm0_extraOffset', 'm1', 'm2', 'm2_extraOffset',
                                 'numElements', 'flags', 'streamPtr']
  assert sig[0][0] == 'double*' and sig[2][0] == 'double'


def test_generation_errors():
  X = Tensor([4, 4], Addressing.STRIDED, alias='X')
  Y = Tensor([4, 4], Addressing.STRIDED, alias='Y')
  # written through one mapping, read through the transpose: points depend on each other
  d = ElementwiseDescr([ot.assign(ot.tensor(X, indices=[-1, -2]),
                                  ot.tensor(X, indices=[-2, -1]) + 1.0)])
  with pytest.raises(GenerationError):
    _gen([d])
  # a rank-1 destination under a rank-2 index space: every n1 writes the same entry
  V = Tensor([4], Addressing.STRIDED, alias='V')
  d = ElementwiseDescr([ot.assign(ot.tensor(V, indices=[-1]), ot.tensor(Y, indices=[-1, -2]))])
  with pytest.raises(GenerationError):
    _gen([d])
  # statements come after the chain they fuse into, never before it
  A = Tensor([4, 4], Addressing.NONE, alias='A')
  m = MultilinearDescr(SubTensor(Y), [SubTensor(A), SubTensor(X)], [[0, -1], [-1, 1]],
                       [[0, 1], [0, 1]])
  e = ElementwiseDescr([ot.assign(ot.tensor(Y, indices=[-1, -2]), 1.0)])
  with pytest.raises(GenerationError):
    _gen([e, m])
  # sparse operands
  S = Tensor([4, 4], Addressing.STRIDED, spp=ListSPP([(0, 0), (1, 1)], (4, 4)), alias='S')
  d = ElementwiseDescr([ot.assign(ot.tensor(Y, indices=[-1, -2]), ot.tensor(S, indices=[-1, -2]))])
  with pytest.raises(GenerationError):
    _gen([d])
  # scalars are not assignable
  s = Tensor([], Addressing.SCALAR, alias='s')
  with pytest.raises(GenerationError):
    _gen([ElementwiseDescr([ot.assign(ot.scalar(s), 1.0)])])
  with pytest.raises(GenerationError):
    ot.tensor(X, indices=[-1])


def test_disjoint_columns_of_one_tensor_are_allowed():
  """Writing several columns of one tensor through bounding boxes is fine: no entry is reached
  twice (the plastic return map does this)."""
  g = catalog.get('pw_plastic_f64').generator
  g.generate()
  assert g.get_emitted().schedule['stores'] == 6


def test_plan_and_phases():
  assert pw.plan(504) == (2, 4)       # 1008 of 1024 lane-points busy
  assert pw.plan(56) == (18, 4)
  assert pw.plan(5000) == (1, 4)
  g = catalog.get('pw_update_f64').generator
  g.generate()
  src = g.get_kernel()
  assert src.index('if (ok_3) c0_3') < src.index('if (ok_0) {') < src.index('if (ok_0 && w')
  em = g.get_emitted()
  assert em.bytes_per_element == 504 * 8 * 3 and em.elements_per_block == 2
  assert 'for (unsigned t0' not in src
  big = Tensor([100, 50], Addressing.STRIDED, alias='B')
  g2 = _gen([ElementwiseDescr([ot.assign(ot.tensor(big, indices=[-1, -2]),
                                         ot.tensor(big, indices=[-1, -2]) * 2.0)])])
  assert 'for (unsigned t0 = 0; t0 < 5000u; t0 += 1024u)' in g2.get_kernel()


def test_literals_take_the_kernel_type():
  X, x = _x()
  g = _gen([ElementwiseDescr([ot.assign(x, x * 0.1)])], FloatingPointType.FLOAT)
  assert '0.1f' in g.get_kernel()
  X, x = _x()
  g = _gen([ElementwiseDescr([ot.assign(x, x * 0.1)])], F64)
  assert '0.1f' not in g.get_kernel() and '* 0.1)' in g.get_kernel()


# ---- the oracle, pinned ----
def _inputs(entry, n, seed=0, flags=None, offsets=None):
  rng = np.random.default_rng(seed)
  arrays = {}
  for key, t in entry.tensors.items():
    if t.addressing == Addressing.SCALAR:
      arrays[t] = 1.25
    elif t.data is None:
      vol = t.get_real_volume()
      size = n * vol + (offsets or {}).get(t, 0)
      arrays[t] = rng.uniform(-1, 1, size=size)
  return OracleInputs(arrays, n, flags=flags, offsets=offsets)


def test_update_closed_form_flags_offsets():
  e = catalog.get('pw_update_f64')
  Q, I, dt = e.tensors['Q'], e.tensors['I'], e.tensors['dt']
  n = 5
  inp = _inputs(e, n, offsets={Q: 3, I: 7})
  inp.flags = np.array([1, 0, 1, 1, 0])
  q0, i0 = inp.arrays[Q].copy(), inp.arrays[I].copy()
  out = pw_run(e.descrs, inp)[Q]
  for el in range(n):
    sq = slice(3 + el * 504, 3 + (el + 1) * 504)
    si = slice(7 + el * 504, 7 + (el + 1) * 504)
    want = q0[sq] + 1.25 * i0[si] if inp.flags[el] else q0[sq]
    np.testing.assert_array_equal(out[sq], want)
  np.testing.assert_array_equal(out[:3], q0[:3])


def _invariants(q):
  """Mean stress and sqrt(J2) per node of a (n, 9, 56) view... given as (elements, 56, 9)."""
  s = q[..., :6]
  m = s[..., :3].mean(axis=-1)
  dev = s[..., :3] - m[..., None]
  j2 = 0.5 * (dev ** 2).sum(-1) + (s[..., 3:] ** 2).sum(-1)
  return m, np.sqrt(j2)


def test_plastic_return_map_invariants():
  e = catalog.get('pw_plastic_f64')
  Q = e.tensors['Q']
  n = 40
  inp = _inputs(e, n, seed=4)
  inp.arrays[e.tensors['cohesion']] = 1.2
  inp.arrays[e.tensors['tanphi']] = 0.2
  q0 = inp.arrays[Q].copy()
  out = pw_run(e.descrs, inp)[Q].copy()
  before = q0.reshape(n, 9, 56).transpose(0, 2, 1)
  after = out.reshape(n, 9, 56).transpose(0, 2, 1)
  m0, tau0 = _invariants(before)
  m1, tau1 = _invariants(after)
  yld = 1.2 - m0 * 0.2
  yielded = tau0 > yld
  assert 0.1 < yielded.mean() < 0.9
  np.testing.assert_allclose(m1, m0, rtol=0, atol=1e-14)                 # mean stress kept
  np.testing.assert_allclose(tau1[yielded], yld[yielded], rtol=1e-13)    # on the yield surface
  np.testing.assert_array_equal(after[~yielded], before[~yielded])       # elastic nodes untouched
  np.testing.assert_array_equal(after[..., 6:], before[..., 6:])         # velocities untouched
  # idempotent: a stress on the surface stays there
  inp2 = OracleInputs({Q: out.copy(), e.tensors['cohesion']: 1.2, e.tensors['tanphi']: 0.2}, n)
  again = pw_run(e.descrs, inp2)[Q]
  np.testing.assert_allclose(again, out, rtol=0, atol=1e-14)


def test_types_known_answers():
  """Hand-computed values at two points: C++ integer division/remainder, the while loop, the
  ternary between a shared and a constant-valued operand, the bounding box and the bitnot."""
  e = catalog.get('t_pw_types')
  T = e.tensors
  n = 1
  A = [np.zeros(24)]
  A[0][0] = 0.3            # point (0, 0)
  A[0][1 + 8 * 2] = -0.7   # point (1, 2)
  K = np.zeros(24)
  K[0] = -7
  K[1 + 8 * 2] = 7
  Wv = np.arange(8, dtype=np.float64) + 2.0     # W[0] = 2
  inp = OracleInputs({T['A']: A, T['W']: Wv, T['K']: K, T['D']: np.full(50, 99.0),
                      T['M']: np.zeros(24), T['s']: 1.5}, n)
  out = pw_run(e.descrs, inp)
  D, M = out[T['D']], out[T['M']]
  # a = 0.3: x = 0.5; t = 0.1 -> 0.2 -> 0.4 -> 0.8 (3 doublings); ternary takes W[0] = 2
  d00 = 0.8 + 2.0 * 1.5 + 0.3 ** 3 + 1.0 + math.log1p(0.3) + 1.0 + 0.0
  # a = -0.7: x = 0.9; t -> 1.6 (4 doublings); ternary takes Cv[1] = -1; round(-2.8) = -3,
  # floor(-2.1) = -3; a > -0.5 is false, so its bitnot negates the result
  d12 = -(1.6 - 1.5 + (-0.5) ** 3 + math.exp(-0.7) + math.log1p(0.7) - 3.0 - 3.0)
  assert D[(1 + 0) + 10 * (1 + 0)] == pytest.approx(d00, rel=1e-15)
  assert D[(1 + 1) + 10 * (1 + 2)] == pytest.approx(d12, rel=1e-15)
  # -7 / 3 = -2 and -7 % 3 = -1 in C++; a > 0; 3 doublings; k > 0 is false
  assert M[0] == -2 - 10 + 100 + 3000
  # 7 / 3 = 2, 7 % 3 = 1; a < 0; 4 doublings; k > 0 and a < 0.5
  assert M[1 + 8 * 2] == 2 + 10 + 4000 + 10000
  # the bounding box [1, 9) x [1, 4) of D is written, its border is not
  border = np.ones((10, 5), dtype=bool)
  border[1:9, 1:4] = False
  assert (D.reshape(5, 10).T[border] == 99.0).all()


def test_math_known_answers():
  e = catalog.get('t_pw_math_f32')
  X, Y = e.tensors['X'], e.tensors['Y']
  xs = np.array([1.0, -0.5, 0.25])
  x = np.zeros(180)
  x[:3] = xs
  out = pw_run(e.descrs, OracleInputs({X: x, Y: np.zeros(180)}, 1))[Y]
  for i, v in enumerate(xs):
    u = abs(v) + 0.5
    want = (math.sin(v) + math.cos(v) * 0.5 + math.tanh(v) + math.sinh(v) * 0.25 + u ** -0.5
            + u ** (1 / 3) + math.atan(v) + math.erf(v) + math.expm1(v * 0.5) + math.log(u)
            + u ** 1.7 + math.sqrt(u) / (u + 1.0) + math.asinh(v) + math.cosh(v) * 0.1)
    assert out[i] == pytest.approx(want, rel=1e-14)


def test_oracle_mutation_is_caught():
  """A plausible mistake (floor division for C++ integer division) changes a known answer."""
  import oracle.pointwise_oracle as po
  v = po._Value(np.array([[-7]]), 'int')
  three = po._Value(np.array([[3]]), 'int')
  r = po._Run.__new__(po._Run)
  r.shape = (1, 1)
  assert r.operation('DIV', [v, three]).data[0, 0] == -2
  assert r.operation('MOD', [v, three]).data[0, 0] == -1
  assert (-7) // 3 == -3     # what a floor-division oracle would have said


def test_empty_batch_and_flags_leave_everything():
  e = catalog.get('pw_update_f64')
  Q = e.tensors['Q']
  inp = _inputs(e, 3, flags=np.zeros(3, dtype=int))
  q0 = inp.arrays[Q].copy()
  np.testing.assert_array_equal(pw_run(e.descrs, inp)[Q], q0)
  inp = _inputs(e, 3)
  inp.num_elements = 0
  q0 = inp.arrays[Q].copy()
  np.testing.assert_array_equal(pw_run(e.descrs, inp)[Q], q0)


# ---- statements fused into a contraction chain ----
def test_fused_statements_generate_one_kernel():
  g = catalog.get('t_fused_gemm_pw').generator
  g.generate()
  em = g.get_emitted()
  assert em.schedule.get('epilogue') == 'pointwise' and em.emitter == 'tuned'
  assert 'fmax(' in g.get_kernel()
  # the scalar of the statements is a launcher argument after the chain's operands
  assert [t for t, _ in g.launcher_signature()][:6] == ['double*', 'unsigned', 'double*',
                                                         'double*', 'unsigned', 'double']


def test_fusion_rules():
  from tensorforge_amd import workloads as W
  w = W.gemm(F64)
  C, B = w.tensors['C'], w.tensors['B']
  other = ot.tensor(B, indices=[-1, -2])
  c = ot.tensor(C, indices=[-1, -2])
  # statements writing another tensor do not fuse (reading one does)
  with pytest.raises(GenerationError):
    _gen(list(w.descrs) + [ElementwiseDescr([ot.assign(other, c + other)])])
  w = W.gemm(F64)
  c = ot.tensor(w.tensors['C'], indices=[-1, -2])
  g = _gen(list(w.descrs) + [ElementwiseDescr([ot.assign(c, c + ot.tensor(w.tensors['B'],
                                                                        indices=[-1, -2]))])])
  assert g.get_emitted().schedule.get('epilogue') == 'pointwise'
  # statements before the chain
  w = W.gemm(F64)
  c = ot.tensor(w.tensors['C'], indices=[-1, -2])
  with pytest.raises(GenerationError):
    _gen([ElementwiseDescr([ot.assign(c, c * 2.0)])] + list(w.descrs))
  # the plain emitter does not fuse
  w = W.gemm(F64)
  c = ot.tensor(w.tensors['C'], indices=[-1, -2])
  g = Generator(list(w.descrs) + [ElementwiseDescr([ot.assign(c, c * 2.0)])],
                Context('gfx950', 'hip', F64), emitter='plain')
  with pytest.raises(GenerationError):
    g.generate()


def test_fused_oracle_closed_form():
  """The oracle runs the chain, then the statements on what it wrote: max(C + A B, 0) * dt."""
  import sys
  import os
  sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
  from harness import make_inputs, oracle_outputs
  e = catalog.get('t_fused_gemm_pw')
  e.generator.register()
  n = 4
  host = make_inputs(e, n, seed=1)
  out = oracle_outputs(e, host, n)
  A = host['A'].reshape(56, 56).T
  for el in range(n):
    B = host['B'][el * 504:(el + 1) * 504].reshape(9, 56).T
    C = host['C'][el * 504:(el + 1) * 504].reshape(9, 56).T
    want = np.maximum(C + A @ B, 0.0) * host['dt']
    got = out['C'][el * 504:(el + 1) * 504].reshape(9, 56).T
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
