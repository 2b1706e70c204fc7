"""Named, prebuilt kernels: the benchmark workloads and the chains the test-suite runs.

``build_catalog()`` generates every entry and compiles each into ``lib/libtf_<entry>.so``;
``load(name)`` rebuilds the entry's descriptions (kernel names are deterministic) and binds the
launcher from the prebuilt object. A missing object is an error, never a fallback.
"""
import concurrent.futures
import os
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

from . import workloads as W
from .common.aux import generate_tmp_matrix, generate_tmp_tensor
from .common.basic_types import Addressing, FloatingPointType
from .common.context import Context
from .common.matrix.boundingbox import BoundingBox
from .common.matrix.spp import ListSPP
from .common.matrix.tensor import SubTensor, Tensor
from .generators.descriptions import GemmDescr, MultilinearDescr
from .generators.generator import Generator
from .runtime import build as rt_build
from .runtime import loader as rt_loader

F64 = FloatingPointType.DOUBLE
F32 = FloatingPointType.FLOAT


@dataclass
class Entry:
  name: str
  descrs: List
  fp_type: FloatingPointType
  tensors: Dict[str, Tensor]
  emitter: str = 'auto'
  schedule: Optional[object] = None
  meta: Dict = field(default_factory=dict)
  generator: Optional[Generator] = None

  def make_generator(self):
    ctx = Context('gfx950', 'hip', self.fp_type)
    self.generator = Generator(self.descrs, ctx, emitter=self.emitter, schedule=self.schedule)
    return self.generator


_REGISTRY: Dict[str, Callable[[], Entry]] = {}


def register(name):
  """Register an entry factory; an entry named ``*_plain`` is always built with the plain
  emitter (the baseline its ``*_tuned`` / ``*_auto`` siblings are compared with)."""
  def deco(fn):
    if name.endswith('_plain'):
      def make(fn=fn):
        entry = fn()
        entry.emitter = 'plain'
        return entry
      _REGISTRY[name] = make
    else:
      _REGISTRY[name] = fn
    return fn
  return deco


def names():
  return sorted(_REGISTRY)


def get(name) -> Entry:
  entry = _REGISTRY[name]()
  entry.name = name
  entry.make_generator()
  return entry


def so_path(name):
  return os.path.join(rt_build.LIB_DIR, f'libtf_{name}.so')


def load(name):
  """``(entry, kernel)`` for a prebuilt entry."""
  entry = get(name)
  entry.generator.generate()
  kernel = rt_loader.load_kernels(so_path(name), [entry.generator])[0]
  return entry, kernel


def _build_one(name):
  entry = get(name)
  entry.generator.generate()
  return rt_build.compile_generators([entry.generator], f'tf_{name}')


def build_catalog(selection=None, jobs=None):
  rt_build.build_runtime()
  todo = list(selection) if selection else names()
  groups = [] if selection else list(_GROUPS)
  jobs = jobs or min(8, os.cpu_count() or 1)
  with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
    futs = [pool.submit(_build_one, n) for n in todo]
    futs += [pool.submit(build_group, g, part) for g in groups for part in ('tuned', 'plain')]
    if not selection:
      futs += [pool.submit(build_omp, (n,)) for n in OMP_ENTRIES]
    return [f.result() for f in futs]


# omptarget (host C++/OpenMP) launchers of the CPU-baseline chains, in lib/libomp_<entry>.so
OMP_ENTRIES = ('ader_f64', 'gemm_f64')


def omp_so_path(name):
  return os.path.join(rt_build.LIB_DIR, f'libomp_{name}.so')


def omp_generator(name):
  """The omptarget backend's generator of catalog chain ``name`` (an ``OMP_ENTRIES`` name: the
  ``*_tuned`` entry's descriptions), generated."""
  entry = _REGISTRY[f'{name}_tuned']()
  gen = Generator(entry.descrs, Context('gfx950', 'omptarget', entry.fp_type))
  gen.generate()
  return entry, gen


def build_omp(names=OMP_ENTRIES):
  return [rt_build.compile_host_source(omp_generator(n)[1].get_source(), f'omp_{n}',
                                       rt_build.LIB_DIR) for n in names]


def load_omp(name):
  """(entry, generator, ctypes function) of a prebuilt omptarget launcher."""
  import ctypes
  entry, gen = omp_generator(name)
  path = omp_so_path(name)
  if not os.path.exists(path):
    raise RuntimeError(f'{path} is missing: build the package first')
  fn = getattr(ctypes.CDLL(path), f'launcher_{gen.get_base_name()}')
  fn.restype = None
  return entry, gen, fn


def omp_arguments(gen, arrays, n, flags=None, offsets=None):
  """ctypes arguments of an omptarget launcher over ``n`` elements, in its parameter order:
  ``arrays`` maps each global tensor's alias (its generated name when it has none) to a
  C-contiguous host array, a list of arrays (a pointer-based tensor: one per element) or a number
  (a floating-point scalar); ``offsets`` the extra offsets by the same keys (default 0); ``flags``
  an optional uint32 array. Returns (arguments, objects to keep alive)."""
  import ctypes
  offsets = offsets or {}
  args, keep = [], []
  for t in gen._global_tensors:
    key = t.alias if t.alias is not None else t.name
    if t.addressing == Addressing.SCALAR:
      ty = gen._arg_type(t)
      if ty not in (F64, F32):
        raise TypeError(f'scalar {key}: only floating-point scalars are passed here')
      v = arrays[key] if t.data is None else \
          (float(t.data[0]) if not isinstance(t.data, dict) else 0.0)
      args.append(ctypes.c_double(v) if ty == F64 else ctypes.c_float(v))
      continue
    if t.has_values() and len(t.get_values()) < 16 and key not in arrays:
      args.append(ctypes.c_void_p(0))        # values baked into the launcher: not read
      if t.addressing != Addressing.NONE:
        args.append(ctypes.c_uint(0))
      continue
    v = arrays[key]
    if isinstance(v, (list, tuple)):
      table = (ctypes.c_void_p * len(v))(*[x.ctypes.data for x in v])
      keep.append(table)
      args.append(ctypes.cast(table, ctypes.c_void_p))
    else:
      args.append(ctypes.c_void_p(v.ctypes.data))
    if t.addressing != Addressing.NONE:
      args.append(ctypes.c_uint(offsets.get(key, 0)))
  args.append(ctypes.c_size_t(n))
  args.append(ctypes.c_void_p(flags.ctypes.data) if flags is not None else ctypes.c_void_p(0))
  stream = ctypes.c_int(0)                 # streamPtr: a stream object, not read on the host
  keep.append(stream)
  args.append(ctypes.cast(ctypes.pointer(stream), ctypes.c_void_p))
  return args, keep


# --------------------------------------------------------------- grouped launches

@dataclass
class GroupEntry:
  """Several chains launched together: ``members`` (tuned, streaming schedule) and the
  grouped launcher over them, plus the same chains with the plain emitter (built into a
  separate object: plain and tuned launchers of a chain share the kernel name)."""
  name: str
  fp_type: object
  workloads: list
  members: List[Entry] = field(default_factory=list)
  plain: List[Entry] = field(default_factory=list)
  grouped: object = None
  meta: dict = field(default_factory=dict)


_GROUPS: Dict[str, Callable[[], GroupEntry]] = {}


def register_group(name):
  def deco(fn):
    _GROUPS[name] = fn
    return fn
  return deco


def group_names():
  return list(_GROUPS)


def get_group(name) -> GroupEntry:
  from .grouped import GroupedGenerator
  g = _GROUPS[name]()
  g.name = name
  from .tune.orders import chosen_schedule
  for w in g.workloads:
    for emitter, dst in (('tuned', g.members), ('plain', g.plain)):
      # tuned members use the autotuned schedule of their order (committed measurements)
      sched = chosen_schedule(w.meta['b']) if emitter == 'tuned' else None
      e = _wl_entry(w, emitter, sched)
      e.name = f'{name}_b{w.meta["b"]}_{emitter}'
      e.make_generator()
      e.generator.generate()
      dst.append(e)
  g.grouped = GroupedGenerator([e.generator for e in g.members])
  g.grouped.generate()
  return g


def group_so_path(name, part):
  return os.path.join(rt_build.LIB_DIR, f'libtf_{name}_{part}.so')


def build_group(name, part):
  g = get_group(name)
  gens = [e.generator for e in g.members] + [g.grouped] if part == 'tuned' else \
      [e.generator for e in g.plain]
  return rt_build.compile_generators(gens, f'tf_{name}_{part}')


def load_group(name):
  """(GroupEntry, tuned member kernels, plain member kernels, grouped kernel)."""
  g = get_group(name)
  tuned = rt_loader.load_kernels(group_so_path(name, 'tuned'),
                                 [e.generator for e in g.members] + [g.grouped])
  plain = rt_loader.load_kernels(group_so_path(name, 'plain'), [e.generator for e in g.plain])
  return g, tuned[:-1], plain, tuned[-1]


@register_group('ader_mixed_f64')
def _mixed_f64():
  return GroupEntry(name='', fp_type=F64, workloads=W.ader_mixed(F64),
                    meta={'bases': list(W.MIXED_BASES)})


# --------------------------------------------------------------- workloads

def _wl_entry(w, emitter, schedule=None):
  return Entry(name='', descrs=w.descrs, fp_type=w.fp_type, tensors=w.tensors, emitter=emitter,
               schedule=schedule, meta=w.meta)


for _emitter in ('plain', 'tuned'):
  for _fp, _fpn in ((F64, 'f64'), (F32, 'f32')):
    register(f'gemm_{_fpn}_{_emitter}')(
      lambda fp=_fp, e=_emitter: _wl_entry(W.gemm(fp), e))
    register(f'ader_{_fpn}_{_emitter}')(
      lambda fp=_fp, e=_emitter: _wl_entry(W.ader_volume(fp), e))
    register(f'ader_dense_{_fpn}_{_emitter}')(
      lambda fp=_fp, e=_emitter: _wl_entry(W.ader_volume(fp, sparse_k=False), e))

# flux integrals either side of the volume chain (reference benchmarks/fused/programs)
for _emitter in ('plain', 'tuned'):
  for _ord in ('mod', 'orig'):
    register(f'flux_local_{_ord}_f64_{_emitter}')(
      lambda e=_emitter, o=_ord: _wl_entry(W.flux_local(F64, o), e))
    register(f'flux_neigh_{_ord}_f64_{_emitter}')(
      lambda e=_emitter, o=_ord: _wl_entry(W.flux_neigh(F64, o), e))

# time-derivative step (reference benchmarks/fused/programs/seissol-deriv_*.cf): aligned (one
# streamed array, per-term contraction ranges 53/54/55) and not aligned (three streamed arrays
# from row 1, 35 destination rows)
for _emitter in ('plain', 'tuned'):
  register(f'deriv_f64_{_emitter}')(
    lambda e=_emitter: _wl_entry(W.ader_derivative(F64, aligned=True), e))
  register(f'deriv_na_f64_{_emitter}')(
    lambda e=_emitter: _wl_entry(W.ader_derivative(F64, aligned=False), e))


# FP64 matrix-core form of the dense chains: the single GEMM, the order-6 volume chain with
# dense K_d (its shared operands fit LDS as doubles in 8-row blocks)
register('gemm_f64_mfma_fp64')(lambda: _wl_entry(W.gemm(F64), 'mfma_fp64'))
register('ader_dense_f64_mfma_fp64')(
  lambda: _wl_entry(W.ader_volume(F64, sparse_k=False), 'mfma_fp64'))
# its other shapes and destination placements (schedule knobs; each built and tested, none the
# default)
MFMA64_OPTIONS = {
  'mfma64w8k8acc': dict(mfma_waves=8, mfma_k_block=8, mfma_dest='acc'),
  'mfma64w8k8end': dict(mfma_waves=8, mfma_k_block=8, mfma_dest='end'),
  'mfma64w4k8regs': dict(mfma_waves=4, mfma_k_block=8, mfma_dest='regs'),
}


def _mfma64_option(kw):
  from .emit.tuned import Schedule
  return _wl_entry(W.gemm(F64), 'mfma_fp64', Schedule(**kw))


for _tag, _kw in MFMA64_OPTIONS.items():
  register(f'opt_gemm_f64_{_tag}')(lambda kw=_kw: _mfma64_option(kw))

# precision study: order 7 (84 basis functions), dense K_d, FP64 inputs and outputs
for _emitter in ('plain', 'tuned', 'mfma_fp32', 'mfma_bf16x3'):
  register(f'ader7_f64_{_emitter}')(
    lambda e=_emitter: _wl_entry(W.ader_volume(F64, b=84, sparse_k=False), e))


# --------------------------------------------------------------- test chains
# Shapes follow the reference's examples (example/*.py) and test suites (tests/gemms/testsuites).

def _t(shape, addr, bbox=None, **kw):
  if bbox is not None:
    bbox = BoundingBox(*bbox)
  return Tensor(shape, addr, bbox, **kw)


def _sub(t, bbox=None):
  return SubTensor(t, BoundingBox(*bbox) if bbox is not None else None)


def _gemm_entry(trans_a, trans_b, alpha, beta, fp=F64, m=24, n=9, k=20, addr_a=Addressing.NONE,
                addr_b=Addressing.STRIDED, addr_c=Addressing.STRIDED, emitter='plain'):
  A = _t([k, m] if trans_a else [m, k], addr_a, alias='A')
  B = _t([n, k] if trans_b else [k, n], addr_b, alias='B')
  C = _t([m, n], addr_c, alias='C')
  d = GemmDescr(trans_a, trans_b, _sub(A), _sub(B), _sub(C), alpha=alpha, beta=beta)
  return Entry('', [d], fp, {'A': A, 'B': B, 'C': C}, emitter=emitter)


for _ta in (False, True):
  for _tb in (False, True):
    register(f't_gemm_{"T" if _ta else "N"}{"T" if _tb else "N"}_plain')(
      lambda ta=_ta, tb=_tb: _gemm_entry(ta, tb, 1.0, 1.0))
register('t_gemm_alpha_beta_plain')(lambda: _gemm_entry(False, False, 1.5, 0.5))
register('t_gemm_overwrite_f32_plain')(lambda: _gemm_entry(False, True, 1.0, 0.0, fp=F32))
register('t_gemm_alpha_runtime_plain')(lambda: _gemm_entry(False, False, 'alpha', 1.0))
register('t_gemm_ptr_plain')(lambda: _gemm_entry(False, False, 1.0, 1.0,
                                                 addr_a=Addressing.PTR_BASED,
                                                 addr_b=Addressing.PTR_BASED,
                                                 addr_c=Addressing.PTR_BASED))
register('t_gemm_strided_a_plain')(lambda: _gemm_entry(True, False, 1.0, 1.0,
                                                       addr_a=Addressing.STRIDED, m=33, n=7, k=31))


def _auto_variants():
  """Every test chain again with the default emitter choice (tuned where it applies)."""
  for name in [n for n in list(_REGISTRY) if n.startswith('t_') and n.endswith('_plain')]:
    base = _REGISTRY[name]

    def make(base=base):
      entry = base()
      entry.emitter = 'auto'
      return entry
    _REGISTRY[name[:-len('_plain')] + '_auto'] = make


@register('t_bbox_plain')
def _bbox_entry():
  """Padded leading dims and sub-block windows (reference notrans-trans-tests.yaml:22-24)."""
  A = _t([72, 40], Addressing.NONE, ([4, 2], [60, 30]), alias='A')
  B = _t([33, 12], Addressing.STRIDED, ([2, 1], [30, 10]), alias='B')
  C = _t([65, 11], Addressing.STRIDED, ([3, 0], [59, 9]), alias='C')
  d = MultilinearDescr(_sub(C, ([4, 1], [60, 10])), [_sub(A, ([4, 2], [60, 30])),
                                                     _sub(B, ([2, 1], [30, 10]))],
                       [[0, -1], [-1, 1]], [[0, 1], [0, 1]], add=True)
  return Entry('', [d], F64, {'A': A, 'B': B, 'C': C})


@register('t_three_index_plain')
def _three_index_entry():
  """C_ijk = A_il · B_ljk (a 3-index contraction of the cost-model workload)."""
  A = _t([12, 10], Addressing.NONE, alias='A')
  B = _t([10, 6, 5], Addressing.STRIDED, alias='B')
  C = _t([12, 6, 5], Addressing.STRIDED, alias='C')
  d = MultilinearDescr(_sub(C), [_sub(A), _sub(B)], [[0, -1], [-1, 1, 2]],
                       [[0, 1], [0, 1, 2]], add=False)
  return Entry('', [d], F64, {'A': A, 'B': B, 'C': C})


@register('t_sparse_dense_plain')
def _sparse_dense_entry():
  """Sparse × dense with a random 25% pattern (reference sparse_dense_gemms/test_loader.py:42)."""
  import numpy as np
  rng = np.random.default_rng(7)
  mask = rng.random((40, 30)) < 0.25
  A = _t([40, 30], Addressing.STRIDED, spp=ListSPP(W.csr_entries(mask), (40, 30)), alias='A')
  B = _t([30, 9], Addressing.STRIDED, alias='B')
  C = _t([40, 9], Addressing.STRIDED, alias='C')
  d = MultilinearDescr(_sub(C), [_sub(A), _sub(B)], [[0, -1], [-1, 1]], [[0, 1], [0, 1]], add=True)
  return Entry('', [d], F64, {'A': A, 'B': B, 'C': C}, meta={'mask': mask})


@register('t_dense_sparse_plain')
def _dense_sparse_entry():
  """Dense × sparse with a random 15% pattern (reference dense_sparse_gemms/test_loader.py)."""
  import numpy as np
  rng = np.random.default_rng(8)
  mask = rng.random((20, 16)) < 0.15
  A = _t([36, 20], Addressing.STRIDED, alias='A')
  B = _t([20, 16], Addressing.NONE, spp=ListSPP(W.csr_entries(mask), (20, 16)), alias='B')
  C = _t([36, 16], Addressing.STRIDED, alias='C')
  d = MultilinearDescr(_sub(C), [_sub(A), _sub(B)], [[0, -1], [-1, 1]], [[0, 1], [0, 1]])
  return Entry('', [d], F64, {'A': A, 'B': B, 'C': C}, meta={'mask': mask})


@register('t_four_matrices_plain')
def _four_matrices_entry():
  """Q += A x ((B x C) x D), all pointer-based (reference example/four_matrices.py:10-34)."""
  q = _t([56, 56], Addressing.PTR_BASED, ([0, 0], [56, 9]), alias='Q')
  a = _t([56, 56], Addressing.PTR_BASED, ([0, 0], [56, 20]), alias='A')
  b = _t([56, 56], Addressing.PTR_BASED, ([0, 0], [20, 56]), alias='B')
  c = _t([56, 9], Addressing.PTR_BASED, ([0, 0], [56, 9]), alias='C')
  d = _t([9, 9], Addressing.PTR_BASED, ([0, 0], [9, 9]), alias='D')
  mq, ma, mb, mc, md = (_sub(x, y) for x, y in ((q, ([0, 0], [56, 9])), (a, ([0, 0], [56, 20])),
                                             (b, ([0, 0], [20, 56])), (c, ([0, 0], [56, 9])),
                                             (d, ([0, 0], [9, 9]))))
  tmp1 = generate_tmp_matrix(mb, mc)
  tmp2 = generate_tmp_matrix(tmp1, md)
  descrs = [GemmDescr(False, False, mb, mc, _sub(tmp1)),
            GemmDescr(False, False, _sub(tmp1), md, _sub(tmp2)),
            GemmDescr(False, False, ma, _sub(tmp2), mq, alpha=1.0, beta=1.0)]
  return Entry('', descrs, F32, {'Q': q, 'A': a, 'B': b, 'C': c, 'D': d})


@register('t_five_multiplies_plain')
def _five_entry():
  """Q = (((A x B) x (C x B))ᵀ-chain x D) (reference example/five_multiplies.py:10-39)."""
  mq = _sub(_t([9, 9], Addressing.STRIDED, alias='Q'))
  ma = _sub(_t([56, 56], Addressing.STRIDED, alias='A'))
  mb = _sub(_t([56, 9], Addressing.STRIDED, alias='B'))
  mc = _sub(_t([56, 56], Addressing.STRIDED, alias='C'))
  md = _sub(_t([9, 9], Addressing.STRIDED, alias='D'))
  tmp0 = generate_tmp_matrix(ma, mb)
  tmp1 = generate_tmp_matrix(mc, mb)
  tmp2 = generate_tmp_matrix(tmp0, tmp1, trans_a=True)
  descrs = [GemmDescr(False, False, ma, mb, _sub(tmp0)),
            GemmDescr(False, False, mc, mb, _sub(tmp1)),
            GemmDescr(True, False, _sub(tmp0), _sub(tmp1), _sub(tmp2)),
            GemmDescr(False, False, _sub(tmp2), md, mq, alpha=1.0, beta=0.0)]
  return Entry('', descrs, F32, {'Q': mq.tensor, 'A': ma.tensor, 'B': mb.tensor, 'C': mc.tensor,
                                 'D': md.tensor})


@register('t_three_matrices_plain')
def _three_matrices_entry():
  """D += A x (B x C), all strided (reference example/three_matrices.py:9-27)."""
  md = _sub(_t([56, 9], Addressing.STRIDED, ([0, 0], [56, 9]), alias='D'), ([0, 0], [56, 9]))
  ma = _sub(_t([56, 56], Addressing.STRIDED, ([0, 0], [56, 56]), alias='A'), ([0, 0], [56, 56]))
  mb = _sub(_t([56, 9], Addressing.STRIDED, ([0, 0], [56, 9]), alias='B'), ([0, 0], [56, 9]))
  mc = _sub(_t([9, 9], Addressing.STRIDED, ([0, 0], [9, 9]), alias='C'), ([0, 0], [9, 9]))
  tmp1 = generate_tmp_matrix(mb, mc)
  descrs = [GemmDescr(False, False, mb, mc, _sub(tmp1)),
            GemmDescr(False, False, ma, _sub(tmp1), md, alpha=1.0, beta=1.0)]
  return Entry('', descrs, F32, {'D': md.tensor, 'A': ma.tensor, 'B': mb.tensor,
                                 'C': mc.tensor})


def _a_atrans_bc(addr_a, bq, bc, addr_c):
  """Q += A x ((A^T x B) x C) over 20-row windows of 56x56 tensors."""
  box = ([0, 0], bq)
  mq = _sub(_t([56, 56], Addressing.STRIDED, box, alias='Q'), box)
  ma = _sub(_t([56, 56], addr_a, box, alias='A'), box)
  mb = _sub(_t([56, 56], Addressing.STRIDED, box, alias='B'), box)
  mc = _sub(_t([56, 9], addr_c, ([0, 0], bc), alias='C'), ([0, 0], bc))
  tmp1 = generate_tmp_matrix(ma, mb, True, False)
  tmp2 = generate_tmp_matrix(tmp1, mc)
  descrs = [GemmDescr(True, False, ma, mb, _sub(tmp1)),
            GemmDescr(False, False, _sub(tmp1), mc, _sub(tmp2)),
            GemmDescr(False, False, ma, _sub(tmp2), mq, alpha=1.0, beta=1.0)]
  return Entry('', descrs, F32, {'Q': mq.tensor, 'A': ma.tensor, 'B': mb.tensor,
                                 'C': mc.tensor})


# reference example/a_atrans_bc.py:27-42 (variant v0: A strided)
register('t_a_atrans_bc_plain')(
  lambda: _a_atrans_bc(Addressing.STRIDED, [20, 9], [9, 9], Addressing.STRIDED))
# reference example/redundant_load_opt.py:23-40 (A and C shared by all elements)
register('t_redundant_load_opt_plain')(
  lambda: _a_atrans_bc(Addressing.NONE, [20, 8], [8, 8], Addressing.NONE))


@register('t_ab_transb_plain')
def _ab_transb_entry():
  """Q += (A x B^T) x B, Q pointer-based, A shared (reference example/ab_transb.py:19-32)."""
  box = ([0, 0], [20, 9])
  mq = _sub(_t([56, 56], Addressing.PTR_BASED, box, alias='Q'), box)
  ma = _sub(_t([56, 56], Addressing.NONE, box, alias='A'), box)
  mb = _sub(_t([56, 56], Addressing.STRIDED, box, alias='B'), box)
  tmp1 = generate_tmp_matrix(ma, mb, trans_a=False, trans_b=True)
  descrs = [GemmDescr(False, True, ma, mb, _sub(tmp1)),
            GemmDescr(False, False, _sub(tmp1), mb, mq, alpha=1.0, beta=1.0)]
  return Entry('', descrs, F32, {'Q': mq.tensor, 'A': ma.tensor, 'B': mb.tensor})


@register('t_abb_trans_plain')
def _abb_trans_entry():
  """Q += (A x B) x B^T, Q pointer-based, A shared (reference example/abb_trans.py:21-34)."""
  box = ([0, 0], [20, 9])
  mq = _sub(_t([56, 56], Addressing.PTR_BASED, box, alias='Q'), box)
  ma = _sub(_t([56, 56], Addressing.NONE, box, alias='A'), box)
  mb = _sub(_t([56, 56], Addressing.STRIDED, ([0, 0], [9, 20]), alias='B'), ([0, 0], [9, 20]))
  tmp1 = generate_tmp_matrix(ma, mb)
  descrs = [GemmDescr(False, False, ma, mb, _sub(tmp1)),
            GemmDescr(False, True, _sub(tmp1), mb, mq, alpha=1.0, beta=1.0)]
  return Entry('', descrs, F32, {'Q': mq.tensor, 'A': ma.tensor, 'B': mb.tensor})


@register('t_mat_vec_plain')
def _mat_vec_entry():
  """C += A x b (reference example/mat_vec.py:9-20)."""
  c = _t([56, 1], Addressing.STRIDED, alias='C')
  a = _t([56, 9], Addressing.STRIDED, alias='A')
  b = _t([9, 1], Addressing.STRIDED, alias='B')
  d = GemmDescr(False, False, _sub(a), _sub(b), _sub(c), alpha=1.0, beta=1.0)
  return Entry('', [d], F32, {'A': a, 'B': b, 'C': c})


@register('t_csa_plain')
def _csa_entry():
  """C = 13 · A x B (reference example/csa.py:29-46)."""
  a = _t([9, 9], Addressing.STRIDED, alias='A')
  b = _t([9, 9], Addressing.STRIDED, alias='B')
  c = _t([9, 9], Addressing.STRIDED, alias='C')
  d = GemmDescr(False, False, _sub(a), _sub(b), _sub(c), alpha=13.0, beta=0.0)
  return Entry('', [d], F32, {'A': a, 'B': b, 'C': c})


@register('t_lane_wide_plain')
def _lane_wide_entry():
  """Leading extents above one wavefront and small ones (order 8 and order 2 bases)."""
  A = _t([120, 120], Addressing.NONE, alias='A')
  B = _t([120, 9], Addressing.STRIDED, alias='B')
  C = _t([120, 9], Addressing.STRIDED, alias='C')
  d = MultilinearDescr(_sub(C), [_sub(A), _sub(B)], [[0, -1], [-1, 1]], [[0, 1], [0, 1]], add=True)
  return Entry('', [d], F64, {'A': A, 'B': B, 'C': C})


_auto_variants()


# ---- pointwise kernels (ElementwiseDescr) ----
for _fp, _fpn in ((F64, 'f64'), (F32, 'f32')):
  register(f'pw_update_{_fpn}')(lambda fp=_fp: _wl_entry(W.pointwise_update(fp), 'auto'))
register('pw_plastic_f64')(lambda: _wl_entry(W.pointwise_plastic(F64), 'auto'))


@register('t_pw_types')
def _pw_types_entry():
  """Pointwise coverage: pointer-based, shared (NONE) and constant-valued operands, a bounding
  box inside a larger tensor with an extra offset, int and bool data, C++ integer division and
  remainder, casts, a ternary, a while loop and floating math, FP64."""
  from .generators import optree as ot
  from .generators.descriptions import ElementwiseDescr
  FT = FloatingPointType
  A = _t([8, 3], Addressing.PTR_BASED, alias='A')
  Wt = _t([8], Addressing.NONE, alias='W')
  Cv = _t([8], Addressing.NONE, alias='Cv', data=[0.5, -1.0, 2.0, 0.0, 1.5, -0.25, 3.0, 1.0])
  K = _t([8, 3], Addressing.STRIDED, alias='K', datatype=FT.INT)
  D = _t([10, 5], Addressing.STRIDED, alias='D')
  M = _t([8, 3], Addressing.STRIDED, alias='M', datatype=FT.INT)
  s = _t([], Addressing.SCALAR, alias='s')
  a = ot.tensor(A, indices=[-1, -2])
  k = ot.tensor(K, indices=[-1, -2])
  w = ot.tensor(Wt, indices=[-1])
  cv = ot.tensor(Cv, indices=[-1])
  d = ot.tensor(SubTensor(D, BoundingBox([1, 1], [9, 4])), indices=[-1, -2])
  m = ot.tensor(M, indices=[-1, -2])
  x, t, n = ot.temp(), ot.temp(), ot.temp(FT.INT)
  stmts = [
    ot.assign(x, abs(a) + 0.2),
    ot.assign(t, 0.1),
    ot.assign(n, 0),
    ot.loop(t < x, [ot.assign(t, t * 2.0), ot.assign(n, n + 1)]),
    ot.assign(d, t + ot.ternary(a > 0.0, w, cv) * ot.scalar(s) + ot.max(a, -0.5) ** 3.0
              + ot.exp(ot.min(a, 0.0)) + ot.logp1(abs(a)) + ot.round(a * 4.0) + ot.floor(a * 3.0)),
    ot.assign(m, k / 3 + (k % 3) * 10 + ot.cast(a > 0.0, FT.INT) * 100 + n * 1000
              + ot.cast((k > 0) & (a < 0.5), FT.INT) * 10000),
    ot.conditional(ot.bitnot(a > -0.5), [ot.assign(d, -d)]),
  ]
  return Entry('', [ElementwiseDescr(stmts)], F64,
               {'A': A, 'W': Wt, 'Cv': Cv, 'K': K, 'D': D, 'M': M, 's': s})


@register('t_pw_math_f32')
def _pw_math_entry():
  """Pointwise floating math in FP32: circular, hyperbolic (distinct from the circular ones),
  inverse, power forms rewritten at construction (x^-0.5, x^(1/3)), erf and expm1."""
  from .generators import optree as ot
  from .generators.descriptions import ElementwiseDescr
  X = _t([20, 9], Addressing.STRIDED, alias='X')
  Y = _t([20, 9], Addressing.STRIDED, alias='Y')
  x = ot.tensor(X, indices=[-1, -2])
  y = ot.tensor(Y, indices=[-1, -2])
  u = ot.temp()
  stmts = [
    ot.assign(u, abs(x) + 0.5),
    ot.assign(y, ot.sin(x) + ot.cos(x) * 0.5 + ot.tanh(x) + ot.sinh(x) * 0.25 + u ** -0.5
              + u ** (1 / 3) + ot.atan(x) + ot.erf(x) + ot.expm1(x * 0.5) + ot.log(u)
              + ot.pow(u, 1.7) + ot.sqrt(u) / (u + 1.0) + ot.asinh(x) + ot.cosh(x) * 0.1),
  ]
  return Entry('', [ElementwiseDescr(stmts)], F32, {'X': X, 'Y': Y})


# ---- schedule options of the streaming emitter (each built and tested, none the default) ----
STREAM_OPTIONS = {
  'k4': dict(stream_k_block=4),
  'la0': dict(stream_lds_lookahead=0),
  'pf2': dict(stream_prefetch=2),
  'w4': dict(stream_waves=4),
  'pair': dict(stream_pair_reads=True),
  'rows': dict(stream_row_loads=True),
  'dma': dict(stream_dma=True),
}


def _option_entry(w, kw):
  from .emit.tuned import Schedule
  return _wl_entry(w, 'tuned', Schedule(**kw))


for _tag, _kw in STREAM_OPTIONS.items():
  register(f'opt_ader_f64_{_tag}')(lambda kw=_kw: _option_entry(W.ader_volume(F64), kw))
  if _tag not in ('dma', 'rows'):      # those two apply to chains that stream through the ring
    register(f'opt_gemm_f64_{_tag}')(lambda kw=_kw: _option_entry(W.gemm(F64), kw))
register('opt_ader_dense_f64_unrolled')(
  lambda: _option_entry(W.ader_volume(F64, sparse_k=False), dict(stream_loop='never')))


@register('t_pw_big')
def _pw_big_entry():
  """Pointwise over more points per element than one workgroup pass covers (120 x 60 = 7,200
  points: the kernel loops over chunks of 1,024), with a pointer-based source."""
  from .generators import optree as ot
  from .generators.descriptions import ElementwiseDescr
  X = _t([120, 60], Addressing.PTR_BASED, alias='X')
  Y = _t([120, 60], Addressing.STRIDED, alias='Y')
  x = ot.tensor(X, indices=[-1, -2])
  y = ot.tensor(Y, indices=[-1, -2])
  return Entry('', [ElementwiseDescr([ot.assign(y, 2.0 * y + ot.sqrt(abs(x)) - x * y)])], F64,
               {'X': X, 'Y': Y})


# ---- contraction chains with fused pointwise statements ----
def _fused(w, statements_of):
  from .generators.descriptions import ElementwiseDescr
  descrs = list(w.descrs) + [ElementwiseDescr(statements_of(w))]
  return Entry('', descrs, w.fp_type, w.tensors, emitter='auto', meta=w.meta)


def _clamp_scale(dest_alias, scalar_alias):
  from .generators import optree as ot

  def statements(w):
    t = w.tensors[dest_alias]
    s = Tensor([], Addressing.SCALAR, alias=scalar_alias)
    w.tensors[scalar_alias] = s
    d = ot.tensor(t, indices=[-1, -2])
    return [ot.assign(d, ot.max(d, 0.0) * ot.scalar(s))]
  return statements


def _soft_limit(dest_alias):
  from .generators import optree as ot

  def statements(w):
    d = ot.tensor(w.tensors[dest_alias], indices=[-1, -2])
    return [ot.conditional(abs(d) > 2.0, [ot.assign(d, 2.0 * ot.tanh(d * 0.5))])]
  return statements


register('t_fused_gemm_pw')(lambda: _fused(W.gemm(F64), _clamp_scale('C', 'dt')))
register('t_fused_ader_pw')(lambda: _fused(W.ader_volume(F64), _soft_limit('Qp')))


@register('pw_clamp_scale_f64')
def _pw_clamp_scale_entry():
  """C = max(C, 0) * dt alone (the statements of t_fused_gemm_pw), for timing the fusion."""
  from .generators import optree as ot
  from .generators.descriptions import ElementwiseDescr
  C = _t([56, 9], Addressing.STRIDED, alias='C')
  dt = _t([], Addressing.SCALAR, alias='dt')
  c = ot.tensor(C, indices=[-1, -2])
  return Entry('', [ElementwiseDescr([ot.assign(c, ot.max(c, 0.0) * ot.scalar(dt))])], F64,
               {'C': C, 'dt': dt})


def _time_update(dest_alias, source_alias, scalar_alias):
  from .generators import optree as ot

  def statements(w):
    s = Tensor([], Addressing.SCALAR, alias=scalar_alias)
    w.tensors[scalar_alias] = s
    d = ot.tensor(w.tensors[dest_alias], indices=[-1, -2])
    q = ot.tensor(w.tensors[source_alias], indices=[-1, -2])
    return [ot.assign(d, q + ot.scalar(s) * d)]
  return statements


register('t_fused_ader_update')(lambda: _fused(W.ader_volume(F64), _time_update('Qp', 'Q', 'dt')))
