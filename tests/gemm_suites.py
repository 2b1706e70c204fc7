"""The reference's dense GEMM test suites, restated as data, and their kernels.

Cases come from ``tests/gemms/testsuites/*.yaml`` of the reference: per suite, the matrix shapes,
bounding boxes ``[row0, col0, row1, col1]`` (upper bounds exclusive), transposes, alpha, beta and
element count; list-valued parameters expand to their cartesian product, as the reference's
``tests/gemms/test_loader.py:44-52`` does (its ``gemm_type`` parameter selects between two
reference code paths and has no counterpart here). The reference's driver
(``tests/gemms/generate.py:83-140``) checks ``C_sub = alpha op(A_sub) op(B_sub) + beta C_sub``
with every matrix addressed from its full-matrix origin, the sub-block at
``offset = rows * col0 + row0``; the launchers here address the bounding box's corner, so the
tests pass that offset as ``extraOffset``.

The sparse-operand suites and the scale-add suites of the reference's tests are restated the
same way (``SPARSE_SUITES``, ``CSA_SUITES``).

    python tests/gemm_suites.py <out_dir>     (build every case's kernels, plain and auto)
"""
import concurrent.futures
import itertools
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)

BUILD_DIR = os.path.join(ROOT, 'tests', 'gemm_suites_build', '_build')
EMITTERS = ('plain', 'auto')

_SQ = (56, 56)
SUITES = [
  # notrans-notrans-tests.yaml:1-35
  dict(name='LeadingDims_NT_NT', a=_SQ, b=_SQ, c=_SQ, trans_a=False, trans_b=False,
       bbox_a=[[0, 0, 33, 9], [22, 46, 55, 55], [0, 46, 33, 55], [22, 0, 55, 9], [11, 23, 44, 32]],
       bbox_b=[[0, 0, 9, 11], [46, 44, 55, 55], [0, 44, 9, 55], [46, 0, 55, 11], [23, 22, 32, 33]],
       bbox_c=[[0, 0, 33, 11], [0, 44, 33, 55], [22, 0, 55, 11], [22, 44, 55, 55], [11, 22, 44, 33]]),
  # notrans-trans-tests.yaml:1-33
  dict(name='LeadingDims_NT_T', a=(72, 43), b=(65, 59), c=(37, 39), trans_a=False, trans_b=True,
       bbox_a=[[0, 0, 12, 10], [0, 33, 12, 43], [60, 0, 72, 10], [60, 33, 72, 43], [30, 16, 42, 26]],
       bbox_b=[[0, 0, 34, 10], [0, 49, 34, 59], [31, 0, 65, 10], [31, 49, 65, 59], [15, 24, 49, 34]],
       bbox_c=[[0, 0, 12, 34], [0, 5, 12, 39], [25, 0, 37, 34], [25, 5, 37, 39], [12, 2, 24, 36]]),
  # trans-notrans-tests.yaml:1-33
  dict(name='LeadingDims_T_NT', a=(60, 43), b=(39, 41), c=(61, 39), trans_a=True, trans_b=False,
       bbox_a=[[0, 0, 12, 34], [0, 9, 12, 43], [48, 0, 60, 34], [48, 9, 60, 43], [24, 4, 36, 38]],
       bbox_b=[[0, 0, 12, 10], [0, 31, 12, 41], [27, 0, 39, 10], [27, 31, 39, 41], [13, 15, 25, 25]],
       bbox_c=[[0, 0, 34, 10], [0, 9, 34, 19], [27, 0, 61, 10], [27, 9, 61, 19], [13, 4, 47, 14]]),
  # trans-trans-tests.yaml:1-33
  dict(name='LeadingDims_T_T', a=(43, 72), b=(41, 39), c=(39, 61), trans_a=True, trans_b=True,
       bbox_a=[[0, 0, 9, 11], [0, 60, 9, 71], [33, 0, 42, 11], [33, 60, 42, 71], [16, 30, 25, 41]],
       bbox_b=[[0, 0, 33, 9], [0, 29, 33, 38], [7, 0, 40, 9], [7, 29, 40, 38], [3, 14, 36, 23]],
       bbox_c=[[0, 0, 11, 33], [0, 27, 11, 60], [27, 0, 38, 33], [27, 27, 38, 60],
               [13, 13, 24, 46]]),
  # loaders-tests.yaml:1-26 (rows of A: 63, 65)
  dict(name='LoaderA', a=[(63, 43), (65, 43)], b=(39, 41), c=(61, 39), trans_a=True,
       trans_b=False, bbox_a=[[0, 0, 12, 34]], bbox_b=[[0, 0, 12, 10]], bbox_c=[[0, 0, 34, 10]]),
  # loaders-tests.yaml:28-55 (rows of B: 31, 33)
  dict(name='LoaderB', a=_SQ, b=[(31, 56), (33, 56)], c=_SQ, trans_a=False, trans_b=False,
       bbox_a=[[0, 0, 30, 10]], bbox_b=[[0, 0, 10, 12]], bbox_c=[[0, 0, 30, 12]]),
  # transpose.yaml:1-24
  dict(name='Transposition', a=(72, 56), b=_SQ, c=_SQ, trans_a=True, trans_b=False, n=1,
       bbox_a=[[0, 0, 33, 31]], bbox_b=[[0, 0, 33, 9]], bbox_c=[[0, 0, 31, 9]]),
  # alpha-beta-tests.yaml:1-27
  dict(name='AlphaBetaTest', a=(56, 9), b=(9, 9), c=(56, 9), trans_a=False, trans_b=False,
       bbox_a=[[0, 0, 56, 9]], bbox_b=[[0, 0, 9, 9]], bbox_c=[[0, 0, 56, 9]],
       alpha=[1.0, 0.0, 2.0], beta=[1.0, 0.0, 1.5]),
]


# tests/dense_sparse_gemms/testsuites/*.yaml and tests/sparse_dense_gemms/testsuites/*.yaml:
# 56x9 dense times a 9x9 sparse matrix (and 56x9 sparse times 9x9 dense), alpha = beta = 1,
# 100 elements; "random" patterns hold 15 % of the entries in column-major order, "full" all of
# them (tests/dense_sparse_gemms/test_loader.py:20-45). Patterns are drawn with a fixed seed.
SPARSE_SUITES = [
  dict(name='A_B_Dense_Sparse', a=(56, 9), b=(9, 9), trans_a=False, trans_b=False, sparse='b',
       kind='random'),
  dict(name='A_B_Dense_Sparse_TransA', a=(9, 56), b=(9, 9), trans_a=True, trans_b=False,
       sparse='b', kind='random'),
  dict(name='A_B_Sparse_Dense_Full', a=(56, 9), b=(9, 9), trans_a=False, trans_b=False,
       sparse='a', kind='full'),
  dict(name='A_B_Sparse_Dense_Random', a=(56, 9), b=(9, 9), trans_a=False, trans_b=False,
       sparse='a', kind='random'),
  dict(name='A_B_Sparse_Dense_TransB', a=(56, 9), b=(9, 9), trans_a=False, trans_b=True,
       sparse='a', kind='random'),
]

# tests/csa/testsuites/*.yaml: C = alpha A + beta C on a bounding box (alpha-beta-tests.yaml,
# with_zero_patch_test.yaml), as pointwise statements.
CSA_SUITES = [
  dict(name='AlphaBetaTest', a=(9, 10), c=(9, 10), bbox=[0, 0, 9, 10], alpha=[0.5, 1.5],
       beta=[1.5, 12.0], n=3),
  dict(name='ZeroPatch', a=(32, 32), c=(56, 56), bbox=[5, 5, 11, 13], alpha=[2.0], beta=[0.0],
       n=3),
]


def sparse_pattern(shape, kind, seed):
  import random
  if kind == 'full':
    return [(i, j) for j in range(shape[1]) for i in range(shape[0])]
  rng = random.Random(seed)
  count = int(0.15 * shape[0] * shape[1])
  picked = set()
  while len(picked) < count:
    picked.add((rng.randint(0, shape[0] - 1), rng.randint(0, shape[1] - 1)))
  return sorted(picked, key=lambda ij: ij[0] + ij[1] * shape[0])


def sparse_cases():
  return [dict(s, n=100, pattern=sparse_pattern(s[s['sparse']], s['kind'], 7 + i))
          for i, s in enumerate(SPARSE_SUITES)]


def sparse_descriptions(case):
  from tensorforge_amd.common.basic_types import Addressing
  from tensorforge_amd.common.matrix.spp import ListSPP
  from tensorforge_amd.common.matrix.tensor import SubTensor, Tensor
  from tensorforge_amd.generators.descriptions import GemmDescr
  def make(key):
    shape = list(case[key])
    spp = ListSPP(case['pattern'], shape) if case['sparse'] == key else None
    return Tensor(shape, Addressing.STRIDED, alias=key.upper(), spp=spp)
  A, B = make('a'), make('b')
  m = case['a'][1] if case['trans_a'] else case['a'][0]
  n = case['b'][0] if case['trans_b'] else case['b'][1]
  C = Tensor([m, n], Addressing.STRIDED, alias='C')
  d = GemmDescr(case['trans_a'], case['trans_b'], SubTensor(A), SubTensor(B), SubTensor(C),
                alpha=1.0, beta=1.0)
  return [d], {'A': A, 'B': B, 'C': C}


def csa_cases():
  out = []
  for s in CSA_SUITES:
    for alpha, beta in itertools.product(s['alpha'], s['beta']):
      out.append(dict(name=f'{s["name"]}_{len(out)}', a=s['a'], c=s['c'], bbox=s['bbox'],
                      alpha=alpha, beta=beta, n=s['n']))
  return out


def csa_descriptions(case):
  from tensorforge_amd.common.basic_types import Addressing
  from tensorforge_amd.common.matrix.boundingbox import BoundingBox
  from tensorforge_amd.common.matrix.tensor import SubTensor, Tensor
  from tensorforge_amd.generators import optree as ot
  from tensorforge_amd.generators.descriptions import ElementwiseDescr
  r0, c0, r1, c1 = case['bbox']
  box = BoundingBox([r0, c0], [r1, c1])
  A = Tensor(list(case['a']), Addressing.STRIDED, alias='A')
  C = Tensor(list(case['c']), Addressing.STRIDED, alias='C')
  a = ot.tensor(SubTensor(A, box), indices=[-1, -2])
  c = ot.tensor(SubTensor(C, box), indices=[-1, -2])
  stmt = ot.assign(c, case['alpha'] * a + case['beta'] * c)
  return [ElementwiseDescr([stmt])], {'A': A, 'C': C}


def _generate(descrs, emitter):
  from tensorforge_amd.common.basic_types import FloatingPointType
  from tensorforge_amd.common.context import Context
  from tensorforge_amd.generators.generator import Generator
  g = Generator(descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE), emitter=emitter)
  g.generate()
  return g


def sparse_generator(case, emitter):
  descrs, tensors = sparse_descriptions(case)
  return _generate(descrs, emitter), descrs, tensors


def csa_generator(case):
  descrs, tensors = csa_descriptions(case)
  return _generate(descrs, 'auto'), descrs, tensors


def _as_list(v, scalar_types):
  return v if isinstance(v, list) and (not v or not isinstance(v[0], scalar_types)) else [v]


def cases():
  """Every case: ``dict(name, a, b, c, bbox_a, bbox_b, bbox_c, trans_a, trans_b, alpha, beta,
  n)`` with ``a``/``b``/``c`` the (rows, cols) of the stored matrices."""
  out = []
  for s in SUITES:
    shapes = [s[k] if isinstance(s[k], list) else [s[k]] for k in ('a', 'b', 'c')]
    alphas = s.get('alpha', 1.1)
    betas = s.get('beta', 1.1)
    alphas = alphas if isinstance(alphas, list) else [alphas]
    betas = betas if isinstance(betas, list) else [betas]
    for a, b, c, ba, bb, bc, alpha, beta in itertools.product(
        shapes[0], shapes[1], shapes[2], s['bbox_a'], s['bbox_b'], s['bbox_c'], alphas, betas):
      out.append(dict(name=f'{s["name"]}_{len(out)}', a=a, b=b, c=c, bbox_a=ba, bbox_b=bb,
                      bbox_c=bc, trans_a=s['trans_a'], trans_b=s['trans_b'], alpha=alpha,
                      beta=beta, n=s.get('n', 10)))
  return out


def offset(shape, bbox):
  """Offset of the bounding box's corner from the full matrix's origin (column-major)."""
  return shape[0] * bbox[1] + bbox[0]


def descriptions(case):
  """(GemmDescr list, {alias: Tensor}) of one case."""
  from tensorforge_amd.common.basic_types import Addressing
  from tensorforge_amd.common.matrix.boundingbox import BoundingBox
  from tensorforge_amd.common.matrix.tensor import SubTensor, Tensor
  from tensorforge_amd.generators.descriptions import GemmDescr
  tensors, subs = {}, {}
  for key in ('a', 'b', 'c'):
    r0, c0, r1, c1 = case[f'bbox_{key}']
    box = BoundingBox([r0, c0], [r1, c1])
    t = Tensor(list(case[key]), Addressing.STRIDED, box, alias=key.upper())
    tensors[key.upper()] = t
    subs[key] = SubTensor(t, box)
  d = GemmDescr(case['trans_a'], case['trans_b'], subs['a'], subs['b'], subs['c'],
                alpha=case['alpha'], beta=case['beta'])
  return [d], tensors


def generator(case, emitter):
  from tensorforge_amd.common.basic_types import FloatingPointType
  from tensorforge_amd.common.context import Context
  from tensorforge_amd.generators.generator import Generator
  descrs, tensors = descriptions(case)
  g = Generator(descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE), emitter=emitter)
  g.generate()
  return g, descrs, tensors


def so_path(emitter, part, directory=BUILD_DIR):
  return os.path.join(directory, f'libgemm_suites_{emitter}_{part}.so')


PARTS = 8


def part_of(index):
  return index % PARTS


def build(directory=BUILD_DIR, jobs=8):
  from tensorforge_amd.runtime.build import compile_source
  from tensorforge_amd.emit.launcher import translation_unit
  os.makedirs(directory, exist_ok=True)
  all_cases = cases()
  work = []
  for emitter in EMITTERS:
    for part in range(PARTS):
      gens = [generator(c, emitter)[0] for i, c in enumerate(all_cases) if part_of(i) == part]
      work.append((translation_unit(gens), f'gemm_suites_{emitter}_{part}'))
  # these chains differ only in which operand is sparse and in its pattern, which the launcher
  # name does not hash: one object each
  for emitter in EMITTERS:
    for i, c in enumerate(sparse_cases()):
      work.append((translation_unit([sparse_generator(c, emitter)[0]]),
                   f'sparse_suites_{emitter}_{i}'))
  work.append((translation_unit([csa_generator(c)[0] for c in csa_cases()]), 'csa_suites'))
  with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
    list(pool.map(lambda w: compile_source(w[0], w[1], directory, keep_source=False), work))
  return len(all_cases) + len(sparse_cases()) * len(EMITTERS) + len(csa_cases())


if __name__ == '__main__':
  print(build(sys.argv[1] if len(sys.argv) > 1 else BUILD_DIR), 'cases')
