"""Description API, naming, prototypes and error behaviour (CPU)."""
import hashlib

import pytest

from tensorforge_amd.common.aux import generate_tmp_matrix, generate_tmp_tensor
from tensorforge_amd.common.basic_types import Addressing, FloatingPointType
from tensorforge_amd.common.context import Context, Options
from tensorforge_amd.common.exceptions import GenerationError
from tensorforge_amd.common.matrix.boundingbox import BoundingBox
from tensorforge_amd.common.matrix.spp import FullSPP, ListSPP, MaskSPP
from tensorforge_amd.common.matrix.tensor import SubTensor, Tensor
from tensorforge_amd.generators.descriptions import GemmDescr, MultilinearDescr
from tensorforge_amd.generators.generator import Generator
from tensorforge_amd import workloads as W


def ctx(fp=FloatingPointType.DOUBLE):
  return Context('gfx950', 'hip', fp)


def test_alias_package_shares_classes():
  import tensorforge
  from tensorforge.common.context import Context as AliasContext
  from tensorforge.generators.descriptions import MultilinearDescr as AliasDescr
  assert AliasContext is Context and AliasDescr is MultilinearDescr
  assert tensorforge.use_fusedgemm_cost() is True
  assert tensorforge.get_version().strip()


def test_only_gfx950_hip():
  with pytest.raises(ValueError):
    Context('sm_60', 'cuda', FloatingPointType.FLOAT)
  with pytest.raises(ValueError):
    Context('gfx942', 'hip', FloatingPointType.FLOAT)
  with pytest.raises(ValueError):
    Context('gfx950', 'sycl', FloatingPointType.FLOAT)
  # the host backend: the same descriptions as C++/OpenMP launchers (CPU baseline, hosts)
  assert Context('gfx950', 'omptarget', FloatingPointType.FLOAT).backend == 'omptarget'


def test_align_follows_wavefront_words():
  # 64 lanes x 4-byte words: 32 doubles or 64 floats (reference context.py:38-44)
  assert ctx().align(56) == 64 and ctx().align(84) == 96
  assert ctx(FloatingPointType.FLOAT).align(56) == 64


def test_thread_policy_values():
  # lanes per element and elements per block of the reference heuristic
  for b, lanes in ((56, 32), (84, 32), (4, 4), (10, 16), (20, 32), (35, 32)):
    w = W.gemm(b=b)
    g = Generator(w.descrs, ctx(), emitter='plain')
    g.generate()
    assert g.get_emitted().block[0] == lanes, b
  w = W.gemm()
  g = Generator(w.descrs, ctx(), emitter='plain')
  g.generate()
  assert g.get_emitted().block == (32, 8, 1)


def test_bbox_checks():
  with pytest.raises(GenerationError):
    Tensor([4, 4], Addressing.STRIDED, BoundingBox([0, 0], [5, 4]))
  with pytest.raises(GenerationError):
    Tensor([4, 4], Addressing.STRIDED, BoundingBox([2, 0], [6, 3]))


def test_strict_match_mismatch_raises():
  w = W.gemm()
  with pytest.raises(RuntimeError):
    Generator(w.descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE,
                                Options(exact_contraction_length=True)))


def test_operand_naming_and_md5_name():
  w = W.ader_volume()
  g = Generator(w.descrs, ctx())
  g.register()
  t = w.tensors
  # first appearance over [dest] + ops: T1 Q S1 | Q' K1 T1 | T2 Q S2 | ...
  assert [t[k].name for k in ('Q', 'S1', 'Qp', 'K1', 'S2', 'K2', 'S3', 'K3')] == \
    ['m0', 'm1', 'm2', 'm3', 'm4', 'm5', 'm6', 'm7']
  long_name = [x.gen_descr() for x in g._global_tensors] + [str(d) for d in w.descrs]
  digest = hashlib.md5(', '.join(long_name).encode()).hexdigest()[:10]
  assert g.get_base_name() == f'kernel_{digest}'
  assert t['Q'].gen_descr() == 'm0 56×9(56×9) {0..56}×{0..9} strided'
  assert str(w.descrs[1]).startswith('m2 56×9(56×9) {0..56}×{0..9} strided({0..56}×{0..9})[0, 1] = ')


def test_launcher_prototype_and_header():
  w = W.gemm()
  g = Generator(w.descrs, ctx())
  g.generate()
  name = g.get_base_name()
  assert g.get_header() == (
    f'void launcher_{name}(double* m0, unsigned m0_extraOffset, const double* m1, '
    'const double* m2, unsigned m2_extraOffset, size_t numElements, unsigned* flags = nullptr, '
    'void* streamPtr = nullptr);\n')
  assert f'extern "C" void launcher_{name}(' in g.get_launcher()
  assert g.default_generate_call_site().startswith(f'launcher_{name}(')


def test_generate_call_site_uses_aliases():
  w = W.gemm()
  g = Generator(w.descrs, ctx())
  g.register()
  site = g.generate_call_site({'A': 'a', 'B': 'b', 'C': 'c'}, {'B': 'ob', 'C': 'oc'}, 'n', 'f', 's')
  assert site == f'launcher_{g.get_base_name()}(c, oc, a, b, ob, n, f, s);'


def test_scalar_alpha_is_a_by_value_argument():
  a = Tensor([4, 4], Addressing.STRIDED)
  b = Tensor([4, 4], Addressing.STRIDED)
  c = Tensor([4, 4], Addressing.STRIDED)
  g = Generator([GemmDescr(False, False, SubTensor(a), SubTensor(b), SubTensor(c),
                           alpha='alpha')], ctx(FloatingPointType.FLOAT))
  g.generate()
  assert ', float m3, size_t numElements' in g.get_header()


def test_gemm_descr_transpose_maps_storage_dims():
  from tensorforge_amd.emit import ir
  a = Tensor([7, 5], Addressing.NONE)      # stored k x m, used transposed
  b = Tensor([7, 3], Addressing.STRIDED)
  c = Tensor([5, 3], Addressing.STRIDED)
  g = Generator([GemmDescr(True, False, SubTensor(a), SubTensor(b), SubTensor(c))], ctx())
  g.generate()
  p = g.get_program().products[0]
  assert p.n_ranges == [(0, 5), (0, 3)] and p.k_ranges == [(0, 7)]
  assert p.factors[0].labels == [('k', 0), ('n', 0)]
  assert not p.add


def test_tmp_shapes():
  a = SubTensor(Tensor([56, 56], Addressing.NONE, BoundingBox([0, 0], [20, 9])),
                BoundingBox([0, 0], [20, 9]))
  b = SubTensor(Tensor([56, 9], Addressing.STRIDED))
  assert generate_tmp_matrix(a, b, trans_a=True).shape == (9, 9)
  assert generate_tmp_tensor([a, b], [[0, -1], [-1, 1]]).shape == (20, 9)


def test_sparsity_patterns_packed_order():
  lst = ListSPP([(2, 0), (0, 1), (1, 1)], (3, 2))
  assert lst.linear_index((0, 1)) == 1 and lst.linear_index((1, 0)) is None
  assert lst.count_nz() == 3 and lst.nonzeros() == [(2, 0), (0, 1), (1, 1)]
  mask = MaskSPP([[True, False], [True, True]])
  assert mask.linear_index((1, 0)) == 1 and mask.linear_index((1, 1)) == 2
  assert FullSPP((2, 3)).nonzeros()[:3] == [(0, 0), (1, 0), (0, 1)]
  assert ListSPP([(0, 0), (2, 1)]).indexmask.shape == (3, 2)


def test_sparse_destination_rejected():
  a = Tensor([4, 4], Addressing.NONE)
  b = Tensor([4, 4], Addressing.STRIDED)
  c = Tensor([4, 4], Addressing.STRIDED, spp=ListSPP([(0, 0)], (4, 4)))
  g = Generator([MultilinearDescr(SubTensor(c), [SubTensor(a), SubTensor(b)],
                                  [[0, -1], [-1, 1]], [[0, 1], [0, 1]])], ctx())
  with pytest.raises(GenerationError):
    g.generate()


def test_traffic_figures_of_workloads():
  g = Generator(W.gemm().descrs, ctx())
  g.generate()
  assert g.get_emitted().bytes_per_element == 12096
  g = Generator(W.ader_volume().descrs, ctx())
  g.generate()
  assert g.get_emitted().bytes_per_element == 8640
  g = Generator(W.ader_volume(FloatingPointType.FLOAT).descrs, ctx(FloatingPointType.FLOAT))
  g.generate()
  assert g.get_emitted().bytes_per_element == 4320


def test_yateto_frontend_dictionary_gateway():
  """add_tensor / add_operation / generate with yateto-like records (reference
  frontend/yateto.py:214-253, 41-87, 151-170)."""
  import tensorforge
  fe_cls = tensorforge.get_routine_generator(None)

  class Arch:
    name, backend, typename = 'gfx950', 'hip', 'double'

  fe = fe_cls(Arch())
  for name, shape, addr, kind in (('A', [8, 8], '&', 'full'), ('B', [8, 3], 'n*N+o&', 'full'),
                                  ('C', [8, 3], 'n&+o&', 'full')):
    fe.add_tensor({'name': name, 'datatype': 'f64', 'addressing': addr,
                   'storage': {'shape': shape, 'type': kind}, 'values': None,
                   'flags': {'temporary': False, 'constant': False}})
  ref = lambda n: {'name': n, 'spp': None}
  fe.add_operation({'type': 'multilinear', 'result': ref('C'), 'args': [ref('A'), ref('B')],
                    'condition': [], 'target': [[0, -1], [-1, 1]], 'permute': [[0, 1], [0, 1]],
                    'linear': {'alpha': None, 'add': True}})
  lines, routines = [], {}

  class Cache:
    def addRoutine(self, name, writer):
      routines[name] = writer

  fe.generate(lines.append, Cache())
  assert len(routines) == 1
  name = next(iter(routines))
  assert lines[0] == (f'launcher_{name}(C, extraOffset_C, const_cast<const double*>(A), '
                      'const_cast<const double*>(B), extraOffset_B, numElements, flags, streamPtr);')


def test_yateto_frontend_pointwise_statements():
  """add_scalar binds optree statements to yateto tensors and scalars through assignTensor
  (reference frontend/yateto.py:127-149); a chain and pointwise statements become two kernels,
  called in order."""
  import tensorforge
  from tensorforge.generators import optree as ot
  fe_cls = tensorforge.get_routine_generator(None)

  class Arch:
    name, backend, typename = 'gfx950', 'hip', 'double'

  class Spp:
    def __init__(self, shape):
      self.shape = shape

    def nnzbounds(self):
      return [(0, s - 1) for s in self.shape]

  class YTensor:
    def __init__(self, name, shape):
      self._name, self._spp = name, Spp(shape)

    def name(self):
      return self._name

    def eqspp(self):
      return self._spp

  class YScalar:
    datatype = None

    def __init__(self, name):
      self._name = name

    def name(self):
      return self._name

  fe = fe_cls(Arch())
  for name, shape, addr in (('A', [8, 8], '&'), ('B', [8, 3], 'n*N+o&'), ('C', [8, 3], 'n*N+o&')):
    fe.add_tensor({'name': name, 'datatype': 'f64', 'addressing': addr,
                   'storage': {'shape': shape, 'type': 'full'}, 'values': None,
                   'flags': {'temporary': False, 'constant': False}})
  ref = lambda n: {'name': n, 'spp': None}
  fe.add_operation({'type': 'multilinear', 'result': ref('C'), 'args': [ref('A'), ref('B')],
                    'condition': [], 'target': [[0, -1], [-1, 1]], 'permute': [[0, 1], [0, 1]],
                    'linear': {'alpha': None, 'add': False}})
  C, dt = YTensor('C', [8, 3]), YScalar('dt')
  c = ot.tensor(C)
  fe.add_scalar([C], [ot.assign(c, ot.max(c, 0.0) * ot.scalar(dt))], [[-1, -2]])
  lines, routines = [], {}

  class Cache:
    def addRoutine(self, name, writer):
      routines[name] = writer

  fe.generate(lines.append, Cache())
  # the statements act on the chain's destination only: one kernel, as the reference fuses them
  assert len(routines) == 1 and len(lines) == 1
  assert lines[0].startswith('launcher_kernel_') and 'const_cast<const double*>(A)' in lines[0]
  assert lines[0].endswith('dt, numElements, flags, streamPtr);')
  fused = list(routines.values())[0]._generator
  assert fused.get_emitted().schedule.get('epilogue') == 'pointwise'
  assert c.indices == [-1, -2] and c.tensor.bbox.sizes() == [8, 3]

  # statements that write another tensor do not fuse: two kernels, called in order
  fe = fe_cls(Arch())
  for name, shape, addr in (('A', [8, 8], '&'), ('B', [8, 3], 'n*N+o&'), ('C', [8, 3], 'n*N+o&')):
    fe.add_tensor({'name': name, 'datatype': 'f64', 'addressing': addr,
                   'storage': {'shape': shape, 'type': 'full'}, 'values': None,
                   'flags': {'temporary': False, 'constant': False}})
  fe.add_operation({'type': 'multilinear', 'result': ref('C'), 'args': [ref('A'), ref('B')],
                    'condition': [], 'target': [[0, -1], [-1, 1]], 'permute': [[0, 1], [0, 1]],
                    'linear': {'alpha': None, 'add': False}})
  B = YTensor('B', [8, 3])
  c2, b2 = ot.tensor(C), ot.tensor(B)
  fe.add_scalar([C, B], [ot.assign(b2, c2 + b2)], [[-1, -2], [-1, -2]])
  lines, routines = [], {}
  fe.generate(lines.append, Cache())
  assert len(routines) == 2 and len(lines) == 2
  assert lines[1].endswith('(C, extraOffset_C, B, extraOffset_B, numElements, flags, '
                           'streamPtr);')
  pointwise = list(routines.values())[1]._generator
  assert pointwise.get_emitted().schedule['kind'] == 'pointwise'


def test_one_object_refuses_two_kernels_under_one_name():
  """The name hashes neither the floating type nor the sparsity pattern: two such chains in one
  translation unit would bind one launcher to the other's kernel, so it is refused."""
  from tensorforge_amd import workloads as W
  from tensorforge_amd.common.exceptions import GenerationError
  from tensorforge_amd.emit.launcher import translation_unit
  sparse = Generator(W.ader_volume(FloatingPointType.DOUBLE).descrs, ctx(FloatingPointType.DOUBLE))
  dense = Generator(W.ader_volume(FloatingPointType.DOUBLE, sparse_k=False).descrs,
                    ctx(FloatingPointType.DOUBLE))
  for g in (sparse, dense):
    g.generate()
  assert sparse.get_base_name() == dense.get_base_name()
  with pytest.raises(GenerationError):
    translation_unit([sparse, dense])
  assert translation_unit([sparse, sparse]).count('extern "C" void launcher_') == 1
