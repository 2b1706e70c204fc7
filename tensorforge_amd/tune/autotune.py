"""Schedule choice for one description list: the cost model's pick, a pruned search or the sweep.

``autotune(descrs, fp_type, batch, method)`` generates every candidate form the chain admits
(``space.applicable``), then:

* ``method='pick'``: the cost model's choice, nothing timed (the model's mean prediction of
  eight networks trained on the committed MI355X timing database, ``data/timings_mi355x.jsonl.gz``,
  in the project's HIP kernels; trained once per process and device);
* ``method='pruned'``: the ``k - 1`` best-predicted candidates plus the library's default choice
  are compiled and timed, the fastest kept;
* ``method='exhaustive'``: every candidate is compiled and timed (the sweep; the fallback and the
  yardstick of the other two).

Forms outside the database the model was trained on (``EXTRA``: the FP64 matrix-core form, for
FP64 chains with dense shared operands) are never picked by the model; the timed methods time
them whenever the chain admits them.

Timing uses synthetic inputs of the chain's shapes at ``batch`` elements (median of up to 25
launches, as the sweep times). Compiled candidates go to one shared object per chain in
``build_dir``; an object already there (same candidate kernel names) is reused, so a search can be
prepared on a host without a GPU and run on the GPU later.
"""
import hashlib
import os
from dataclasses import dataclass, field
from typing import Dict

import numpy as np

from ..common.basic_types import FloatingPointType
from . import features as F
from . import model as M
from . import space

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'timings_mi355x.jsonl.gz')
# the training settings the cross-validation of the committed database measured (7,921 chains:
# 40 epochs reach what 60 do, in 55 s for the five folds of eight networks)
SETTINGS = dict(epochs=40, batch=2048, lr=0.004, ensemble=8, target='gap', cap=0.25)

_MODELS = {}


def _extra():
  """Candidate forms the cost model was not trained on."""
  return [space.Candidate('mfma_fp64', 'mfma_fp64', None)]


@dataclass
class Choice:
  label: str                        # the chosen candidate (``space.candidate_space`` name)
  candidate: object                 # its space.Candidate
  generator: object                 # its generated Generator (kernel named after the chain)
  method: str
  predicted: Dict[str, float] = field(default_factory=dict)   # label -> predicted log-time gap
  timed: Dict[str, float] = field(default_factory=dict)       # label -> seconds (timed methods)


def cost_model(device='cuda:0'):
  """The cost model trained on the committed database (once per process and device)."""
  key = str(device)
  if key not in _MODELS:
    _MODELS[key], _ = M.train(M.load(DATA), device=device, **SETTINGS)
  return _MODELS[key]


def chain_kind(generator):
  """The random-chain family a description list belongs to, as a feature: 'three_index' when an
  operand has more than two dimensions, 'fused' when a temporary is the product of two
  per-element operands (the volume-chain form), else 'linear'."""
  from ..emit import ir
  prog = generator.get_program()
  if any(len(f.tensor.shape) > 2 for p in prog.products for f in p.factors):
    return 'three_index'
  for p in prog.products:
    if p.dest_src == ir.SRC_TMP and len(p.factors) == 2 and \
        all(f.src == ir.SRC_GLOBAL and f.arg.per_element for f in p.factors):
      return 'fused'
  return 'linear'


def candidates(descrs, fp_type, batch, name_prefix):
  """[(Candidate, Generator)] the chain admits, the feature records of the model's candidates
  (the first ``len(records)`` entries) and, after them, the admitted forms of ``_extra()``."""
  cands = space.applicable(descrs, fp_type, name_prefix=name_prefix)
  if not cands:
    raise ValueError('no candidate form admits this description list')
  kind = chain_kind(cands[0][1])
  info = {'kind': kind}
  recs = [{'label': c.label, 'features': F.features(g, c, batch, info)} for c, g in cands]
  extra = space.applicable(descrs, fp_type, cands=_extra(), name_prefix=name_prefix)
  return cands + extra, recs


def _prefix(descrs, fp_type, batch):
  """A kernel-name prefix per (chain, type, batch): candidates of different chains loaded into one
  process never share a symbol. The chain's kernel name alone does not separate chains that
  differ only in a sparsity pattern, so the prefix also hashes the default form's source."""
  from ..common.context import Context
  from ..generators.generator import Generator
  g = Generator(descrs, Context('gfx950', 'hip', fp_type))
  g.generate()
  h = hashlib.md5(f'{g.get_base_name()}/{fp_type}/{batch}/{g.get_kernel()}'.encode())
  return f'at{h.hexdigest()[:10]}'


def library(cands, build_dir, force=False):
  """The shared object holding every candidate of one chain (compiled when missing, or always
  with ``force``): one object serves every method, and can be built ahead of the GPU run."""
  from ..runtime.build import compile_generators
  gens = [g for _, g in cands]
  name = gens[0].get_base_name() + '_' + hashlib.md5(
    '/'.join(g.get_kernel() for g in gens).encode()).hexdigest()[:8]
  so = os.path.join(build_dir, f'lib{name}.so')
  if force or not os.path.exists(so):
    os.makedirs(build_dir, exist_ok=True)
    compile_generators(gens, name, build_dir)
  return so


def prepare(descrs, fp_type, batch, build_dir, force=False):
  """Compile every candidate of the chain into ``build_dir`` (no GPU needed; an existing object
  is kept unless ``force``); returns the object's path."""
  cands, _ = candidates(descrs, fp_type, batch, _prefix(descrs, fp_type, batch))
  return library(cands, build_dir, force)


def _time(cands, idx, batch, build_dir, device, steps):
  """Seconds per launch of the candidates ``idx`` of ``cands``."""
  from ..device import launch_args
  from ..runtime.loader import load_kernels
  from .measure import synthetic_buffers
  from .sweep import _time_candidate
  gens = [g for _, g in cands]
  kernels = load_kernels(library(cands, build_dir), gens)
  bufs = synthetic_buffers(gens[0], batch, device)
  args = launch_args(gens[0], {k: v for k, v in bufs.items() if k != '__keep__'}, batch)
  out = {}
  for i in idx:
    out[cands[i][0].label] = _time_candidate(kernels[i], args, steps)
  del bufs
  return out


def autotune(descrs, fp_type=FloatingPointType.DOUBLE, batch=65536, method='pick', k=3,
             device='cuda:0', build_dir=None, steps=25) -> Choice:
  """The chosen form of ``descrs`` at ``batch`` elements (see the module docstring)."""
  if method not in ('pick', 'pruned', 'exhaustive'):
    raise ValueError(f'unknown method {method!r}')
  cands, recs = candidates(descrs, fp_type, batch, _prefix(descrs, fp_type, batch))
  labels = [r['label'] for r in recs]
  pred = cost_model(device).predict(M.feature_matrix(recs))
  predicted = dict(zip(labels, (float(p) for p in pred)))
  order = np.argsort(pred)
  if method == 'pick':
    i = int(order[0])
    return Choice(labels[i], cands[i][0], cands[i][1], method, predicted)
  if build_dir is None:
    raise ValueError('the timed methods need build_dir for the compiled candidates')
  extra = list(range(len(recs), len(cands)))
  idx = M.pruned(order, labels, k) + extra if method == 'pruned' else list(range(len(cands)))
  timed = _time(cands, idx, batch, build_dir, device, steps)
  best = min(timed, key=timed.get)
  i = [c.label for c, _ in cands].index(best)
  return Choice(best, cands[i][0], cands[i][1], method, predicted, timed)
