"""Candidate schedules of a chain and the random contraction chains the cost model learns from.

Candidates (tile shapes, LDS layout and the threads-per-element schedule):

* the plain emitter (the reference's thread mapping: a row of lanes per element);
* the column-per-lane tuned emitter (one lane per output column) with 1 or 4 wavefronts per
  workgroup, 8- or 16-deep contraction blocks, and with its LDS regions unpadded;
* the streaming schedule with 4 or 8 wavefronts, 4- or 8-row blocks and LDS read-ahead 0 or 2, plus
  16-row blocks, the paired (interleaved) chunk layout of the shared operands in LDS (FP64) and
  the coalesced row loads of the streamed operand;
* with ``precision='mixed'`` also the matrix-core forms (FP32 MFMA and split BF16x3) of FP64
  chains with dense shared operands; they change the result within 1e-5, so the default
  ``'exact'`` space leaves them out.

A candidate that the chain's shape does not admit (the emitter falls back to another form) is
dropped, so every candidate label names the kernel that is actually built.
"""
import itertools
from dataclasses import asdict, dataclass
from typing import Optional

import numpy as np

from ..common.aux import generate_tmp_tensor
from ..common.basic_types import Addressing, FloatingPointType
from ..common.context import Context
from ..common.matrix.spp import ListSPP
from ..common.matrix.tensor import SubTensor, Tensor
from ..emit.tuned import Schedule
from ..generators.descriptions import MultilinearDescr
from ..generators.generator import Generator


@dataclass
class Candidate:
  label: str
  emitter: str
  schedule: Optional[Schedule]

  def knobs(self):
    return asdict(self.schedule) if self.schedule is not None else {}


def candidate_space(precision='exact'):
  out = [Candidate('plain', 'plain', None)]
  for waves, kb in itertools.product((1, 4), (8, 16)):
    out.append(Candidate(f'cpl_w{waves}_k{kb}', 'tuned',
                         Schedule(stream=False, waves_per_block=waves, k_block=kb)))
  out.append(Candidate('cpl_w4_k16_nopad', 'tuned',
                       Schedule(stream=False, waves_per_block=4, k_block=16, lds_ld_pad=0,
                                lds_elem_pad=0)))
  for waves, kb, la in itertools.product((4, 8), (4, 8), (0, 2)):
    out.append(Candidate(f'stream_w{waves}_k{kb}_a{la}', 'tuned',
                         Schedule(stream_waves=waves, stream_k_block=kb, stream_lds_lookahead=la)))
  out.append(Candidate('stream_w8_k16_a2', 'tuned',
                       Schedule(stream_waves=8, stream_k_block=16, stream_lds_lookahead=2)))
  out.append(Candidate('stream_w8_k8_a2_pair', 'tuned',
                       Schedule(stream_waves=8, stream_k_block=8, stream_lds_lookahead=2,
                                stream_pair_reads=True)))
  out.append(Candidate('stream_w8_k8_a2_rows', 'tuned',
                       Schedule(stream_waves=8, stream_k_block=8, stream_lds_lookahead=2,
                                stream_row_loads=True)))
  if precision == 'mixed':
    out.append(Candidate('mfma_fp32', 'mfma_fp32', None))
    out.append(Candidate('mfma_bf16x3', 'mfma_bf16x3', None))
  return out


def _kind(label):
  return label.split('_')[0]


def _admitted(c, em):
  """The emitted kernel is the form the candidate names (no fallback to another form)."""
  if c.emitter == 'plain':
    return em.emitter == 'plain'
  if c.emitter.startswith('mfma'):
    return em.emitter == 'mfma'
  kind = em.schedule.get('kind', 'column') if em.emitter == 'tuned' else em.emitter
  want = 'stream' if _kind(c.label) == 'stream' else 'column'
  if kind != want:
    return False
  sch = c.schedule
  if want == 'stream':
    # options the chain does not take are silently off in the emitter: not a distinct kernel
    if sch.stream_pair_reads and em.schedule.get('pair') is not True:
      return False
    if sch.stream_row_loads and em.schedule.get('row_loads') is not True:
      return False
    if em.schedule.get('k_block') != sch.stream_k_block and not em.schedule.get('loop'):
      return False
  return True


def applicable(descrs, fp_type, cands=None, name_prefix=None, precision='exact'):
  """[(candidate, generated Generator)] for the candidates the chain admits. With
  ``name_prefix`` every candidate's kernel and launcher get their own name
  (``<prefix>_<label>``), so all of them can go into one shared object."""
  ctx = Context('gfx950', 'hip', fp_type)
  out = []
  seen = set()
  for c in cands or candidate_space(precision):
    g = Generator(descrs, ctx, emitter=c.emitter, schedule=c.schedule)
    if name_prefix is not None:
      g.set_kernel_name(f'{name_prefix}_{c.label}')
    try:
      g.generate()
    except Exception:   # shape not admitted by this emitter
      continue
    if not _admitted(c, g.get_emitted()):
      continue
    # knobs the chain's kernel does not depend on give the same kernel twice: keep the first
    body = g.get_kernel().replace(g.get_base_name(), '@')
    if body in seen:
      continue
    seen.add(body)
    out.append((c, g))
  return out


# ------------------------------------------------------------------ random chains

def _shared(rng, rows, cols, sparsity, name):
  if sparsity > 0:
    mask = rng.random((rows, cols)) >= sparsity
    if not mask.any():
      mask[rng.integers(rows), rng.integers(cols)] = True
    entries = [(int(r), int(c)) for c in range(cols) for r in range(rows) if mask[r, c]]
    return Tensor([rows, cols], Addressing.NONE, spp=ListSPP(entries, (rows, cols)), alias=name)
  return Tensor([rows, cols], Addressing.NONE, alias=name)


def random_chain(rng, max_dim=128, max_element_bytes=96 * 1024):
  """One chain of the learned-cost-model workload: 2-5 GEMMs with dimensions in [4, max_dim],
  shared-operand sparsity in {0, .3, .6, .9}, FP32 or FP64, batch log-uniform in [1e3, 1e6];
  a tenth are 3-index contractions C_ijk = A_il B_ljk. Returns (descrs, fp_type, batch, info)."""
  fp = FloatingPointType.DOUBLE if rng.random() < 0.5 else FloatingPointType.FLOAT
  size = 8 if fp == FloatingPointType.DOUBLE else 4
  batch = int(round(10 ** rng.uniform(3, 6)))
  kind = rng.choice(['linear', 'fused', 'three_index'], p=[0.6, 0.3, 0.1])
  dim = lambda: int(rng.integers(4, max_dim + 1))
  spars = lambda: float(rng.choice([0.0, 0.3, 0.6, 0.9]))
  if kind == 'three_index':
    i, l, j, k = dim(), dim(), int(rng.integers(4, 17)), int(rng.integers(4, 17))
    A = _shared(rng, i, l, spars(), 'A')
    B = Tensor([l, j, k], Addressing.STRIDED, alias='B')
    C = Tensor([i, j, k], Addressing.STRIDED, alias='C')
    d = [MultilinearDescr(SubTensor(C), [SubTensor(A), SubTensor(B)], [[0, -1], [-1, 1, 2]],
                          [[0, 1], [0, 1, 2]], add=True)]
    per = (l * j * k + 2 * i * j * k) * size
    info = {'kind': kind, 'dims': [i, l, j, k], 'gemms': 1}
  elif kind == 'linear':
    n = int(rng.integers(2, 6))
    q = int(rng.integers(4, 33)) if rng.random() < 0.8 else dim()
    ms = [dim() for _ in range(n + 1)]
    Bt = Tensor([ms[0], q], Addressing.STRIDED, alias='B')
    prev = SubTensor(Bt)
    d = []
    for s in range(n):
      A = _shared(rng, ms[s + 1], ms[s], spars(), f'A{s + 1}')
      if s == n - 1:
        dest = SubTensor(Tensor([ms[s + 1], q], Addressing.STRIDED, alias='C'))
      else:
        dest = SubTensor(generate_tmp_tensor([SubTensor(A), prev], [[0, -1], [-1, 1]]))
      d.append(MultilinearDescr(dest, [SubTensor(A), prev], [[0, -1], [-1, 1]],
                                [[0, 1], [0, 1]], add=(s == n - 1 and rng.random() < 0.5)))
      prev = dest
    per = (ms[0] * q + 2 * ms[-1] * q) * size
    info = {'kind': kind, 'dims': ms + [q], 'gemms': n}
  else:
    terms = int(rng.integers(1, 3))
    b, q = dim(), int(rng.integers(4, 17))
    Q = Tensor([b, q], Addressing.STRIDED, alias='Q')
    Qp = Tensor([b, q], Addressing.STRIDED, alias='Qp')
    d = []
    for t in range(terms):
      W = Tensor([q, q], Addressing.STRIDED, alias=f'W{t + 1}')
      K = _shared(rng, b, b, spars(), f'K{t + 1}')
      T = generate_tmp_tensor([SubTensor(Q), SubTensor(W)], [[0, -1], [-1, 1]])
      d.append(MultilinearDescr(SubTensor(T), [SubTensor(Q), SubTensor(W)], [[0, -1], [-1, 1]],
                                [[0, 1], [0, 1]]))
      d.append(MultilinearDescr(SubTensor(Qp), [SubTensor(K), SubTensor(T)], [[0, -1], [-1, 1]],
                                [[0, 1], [0, 1]], add=t > 0))
    per = (2 * b * q + terms * q * q) * size
    info = {'kind': kind, 'dims': [b, q, terms], 'gemms': 2 * terms}
  if per > max_element_bytes:
    return None
  batch = max(1000, min(batch, int((1 << 29) // per)))
  info.update({'fp': 'f64' if fp == FloatingPointType.DOUBLE else 'f32', 'batch': batch,
               'element_bytes': per})
  return d, fp, batch, info


def random_chains(count, seed=5):
  rng = np.random.default_rng(seed)
  out = []
  while len(out) < count:
    c = random_chain(rng)
    if c is not None:
      out.append(c)
  return out
