"""Autotuner pieces that need no GPU: candidate space, random chains, features, evaluation."""
import numpy as np

from tensorforge_amd import workloads as W
from tensorforge_amd.common.basic_types import FloatingPointType
from tensorforge_amd.tune import features as F
from tensorforge_amd.tune import model as M
from tensorforge_amd.tune import space


def test_candidates_of_the_single_gemm_cover_all_emitters():
  w = W.gemm()
  labels = [c.label for c, _ in space.applicable(w.descrs, w.fp_type)]
  assert 'plain' in labels
  assert any(l.startswith('stream') for l in labels)
  assert any(l.startswith('cpl') for l in labels)


def test_random_chains_follow_the_stated_distribution():
  chains = space.random_chains(60, seed=11)
  again = space.random_chains(60, seed=11)
  assert [c[3] for c in chains] == [c[3] for c in again]
  kinds = {c[3]['kind'] for c in chains}
  assert kinds == {'linear', 'fused', 'three_index'}
  for descrs, fp, batch, info in chains:
    assert 1000 <= batch <= 10 ** 6
    assert all(4 <= d <= 128 for d in info['dims'][:2])
    assert fp in (FloatingPointType.FLOAT, FloatingPointType.DOUBLE)
    if info['kind'] == 'linear':
      assert 2 <= info['gemms'] <= 5


def test_feature_vectors_have_the_named_length():
  w = W.ader_volume()
  for c, g in space.applicable(w.descrs, w.fp_type)[:4]:
    v = F.features(g, c, 262144, {'kind': 'fused'})
    assert len(v) == len(F.NAMES) and all(np.isfinite(v))


class _Oracle:
  """A 'model' that predicts the measured times exactly: picks must be the sweep's best."""

  def __init__(self, recs):
    self.t = {tuple(x): np.log(r['time']) for x, r in zip(M.feature_matrix(recs).tolist(), recs)}

  def predict(self, X):
    return np.array([self.t[tuple(x)] for x in X.tolist()])


def test_evaluation_measures_slowdown_against_the_sweep():
  recs = []
  for chain in range(4):
    for j, label in enumerate(['plain', 'cpl_w1_k16', 'stream_w8_k8_a2']):
      recs.append({'chain': chain, 'label': label, 'time': 1.0 + j + chain,
                   'features': [float(chain), float(j)]})
  s = M.evaluate(_Oracle(recs), recs)
  assert s['model_pick']['mean_slowdown'] == 0.0
  assert s['pruned_search']['within_5pct'] == 1.0
  # the default choice is the streaming schedule here: the slowest
  assert s['default']['mean_slowdown'] > 0.5
  train, test = M.split(recs, 0.25, seed=1)
  assert {r['chain'] for r in train}.isdisjoint({r['chain'] for r in test})


def test_ensemble_predicts_the_mean_of_its_networks():
  """The ensemble's score is the mean of its members' scores, so its pick can differ from
  every member's (here member one prefers candidate 0, member two candidate 2, the mean 1)."""
  class Fixed:
    def __init__(self, v):
      self.v = np.array(v, dtype=float)

    def predict(self, X):
      return self.v[:len(X)]
  ens = M.Ensemble([Fixed([0.0, 1.0, 3.0]), Fixed([3.0, 1.0, 0.0])])
  np.testing.assert_allclose(ens.predict(np.zeros((3, 2))), [1.5, 1.0, 1.5])
  recs = [{'chain': 'c', 'label': lab, 'time': t, 'features': [0.0, 0.0]}
          for lab, t in (('plain', 3.0), ('tuned/a', 1.0), ('tuned/b', 2.0))]
  s = M.evaluate(ens, recs, prune_k=1)
  assert s['model_pick']['mean_slowdown'] == 0.0


def test_named_candidates_are_distinct_kernels_of_one_chain():
  """With a name prefix every candidate gets its own kernel and launcher name (one shared object
  holds them all), and knobs the chain's kernel does not depend on give no duplicate."""
  w = W.ader_volume()
  cands = space.applicable(w.descrs, w.fp_type, name_prefix='c7')
  names = [g.get_base_name() for _, g in cands]
  assert len(set(names)) == len(names)
  assert all(n == f'c7_{c.label}' for (c, _), n in zip(cands, names))
  bodies = [g.get_kernel().replace(g.get_base_name(), '@') for _, g in cands]
  assert len(set(bodies)) == len(bodies)
  labels = [c.label for c, _ in cands]
  # the paired LDS layout of the shared operand applies to this FP64 chain
  assert 'stream_w8_k8_a2_pair' in labels


def test_matrix_core_candidates_only_in_the_mixed_precision_space():
  w = W.ader_volume(b=84, sparse_k=False)
  exact = [c.label for c, _ in space.applicable(w.descrs, w.fp_type)]
  mixed = [c.label for c, _ in space.applicable(w.descrs, w.fp_type, precision='mixed')]
  assert not any(l.startswith('mfma') for l in exact)
  assert {'mfma_fp32', 'mfma_bf16x3'} <= set(mixed)


def test_sweep_chain_draws_are_reproducible_per_index():
  from tensorforge_amd.tune import sweep
  a = sweep.chain_descriptions(5, 123)
  b = sweep.chain_descriptions(5, 123)
  c = sweep.chain_descriptions(5, 124)
  assert a[3] == b[3] and a[2] == b[2]
  assert (a[3], a[2]) != (c[3], c[2])


def test_timing_database_size_and_sweep_record():
  """The committed timing database: at least 2,000 timed chains, every candidate timed in two
  passes, counts matching the sweep's record."""
  import json
  import os
  from tensorforge_amd.tune import model as M
  data = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      'tensorforge_amd', 'tune', 'data')
  recs = M.load(os.path.join(data, 'timings_mi355x.jsonl.gz'))
  meta = json.load(open(os.path.join(data, 'timings_mi355x.json')))
  chains = {r['chain'] for r in recs}
  assert len(chains) == meta['timed_chains'] >= 2000
  assert len(recs) == meta['timed_candidates']
  assert all(r['runs'] == 2 for r in recs)
  assert meta['drawn'] >= meta['kept'] >= meta['timed_chains']
  assert sum(p['chains'] for p in meta['measure_passes']) == 2 * meta['timed_chains']


def test_autotune_chain_kind_and_candidate_records():
  """The autotuner's chain family (a feature) and its candidate records, without a GPU: the
  volume chain is 'fused', a GEMM 'linear', a 3-index contraction 'three_index'; every record
  carries the full feature vector; unknown methods are refused."""
  import pytest
  from tensorforge_amd import workloads as W
  from tensorforge_amd.common.basic_types import FloatingPointType
  from tensorforge_amd.tune import autotune as at
  from tensorforge_amd.tune import features as F
  from tensorforge_amd.tune.sweep import chain_descriptions
  from tensorforge_amd.generators.generator import Generator
  from tensorforge_amd.common.context import Context

  def kind(descrs, fp):
    g = Generator(descrs, Context('gfx950', 'hip', fp))
    g.generate()
    return at.chain_kind(g)
  assert kind(W.ader_volume().descrs, FloatingPointType.DOUBLE) == 'fused'
  assert kind(W.gemm().descrs, FloatingPointType.DOUBLE) == 'linear'
  i = next(i for i in range(100) if chain_descriptions(7, i)[3]['kind'] == 'three_index')
  descrs, fp, batch, info = chain_descriptions(7, i)
  assert kind(descrs, fp) == 'three_index'
  w = W.gemm()
  cands, recs = at.candidates(w.descrs, w.fp_type, 65536, 'attest')
  assert len(recs) >= 2 and [c.label for c, _ in cands[len(recs):]] == ['mfma_fp64']
  assert all(len(r['features']) == len(F.NAMES) for r in recs)
  assert len({g.get_emitted().kernel_name for _, g in cands}) == len(cands)
  with pytest.raises(ValueError):
    at.autotune(w.descrs, w.fp_type, 65536, method='fastest')
  # chains that differ only in a sparsity pattern share the reference's kernel name, not the
  # autotuner's symbol prefix
  sp, de = W.ader_volume(), W.ader_volume(sparse_k=False)
  assert at._prefix(sp.descrs, sp.fp_type, 1000) != at._prefix(de.descrs, de.fp_type, 1000)
  # the FP64 matrix-core form comes after the model's candidates, for dense chains only
  cands, recs = at.candidates(de.descrs, de.fp_type, 262144, 'attest2')
  assert [c.label for c, _ in cands[len(recs):]] == ['mfma_fp64']
  cands, recs = at.candidates(sp.descrs, sp.fp_type, 262144, 'attest3')
  assert len(cands) == len(recs)


def test_sweep_pack_averages_passes_and_counts(tmp_path):
  """``sweep.pack``: a candidate timed in two passes keeps the mean of its times and ``runs`` 2,
  features and times are rounded, and the sweep's record sums the directories' counts."""
  import gzip
  import json
  from tensorforge_amd.tune import model as M
  from tensorforge_amd.tune.sweep import pack
  d = tmp_path / 'sw'
  d.mkdir()
  json.dump({'chains': [{'chain': 3}, {'chain': 4}], 'skipped': 5, 'generate_seconds': 2.0,
             'compile_seconds': 8.0}, open(d / 'index.json', 'w'))
  rec = lambda c, lab, t: {'chain': c, 'label': lab, 'time': t, 'batch': 1000, 'info': {},
                           'features': [1.234567891, 2.0]}
  f1, f2 = tmp_path / 'a.jsonl', tmp_path / 'b.jsonl'
  f1.write_text('\n'.join(json.dumps(r) for r in (rec(3, 'plain', 2e-3), rec(3, 'x', 1e-3))))
  f2.write_text('\n'.join(json.dumps(r) for r in (rec(3, 'plain', 4e-3), rec(4, 'x', 5e-4))))
  out = str(tmp_path / 'db.jsonl.gz')
  meta = pack([str(d)], [str(f1), str(f2)], out, 'test')
  recs = M.load(out)
  by = {(r['chain'], r['label']): r for r in recs}
  assert by[(3, 'plain')]['time'] == 3e-3 and by[(3, 'plain')]['runs'] == 2
  assert by[(3, 'x')]['runs'] == 1 and by[(4, 'x')]['time'] == 5e-4
  assert by[(3, 'x')]['features'][0] == 1.234568 and all(r['set'] == 'test' for r in recs)
  assert meta == {'drawn': 7, 'kept': 2, 'generate_cpu_seconds': 2.0, 'compile_cpu_seconds': 8.0,
                  'timed_chains': 2, 'timed_candidates': 3}
  with gzip.open(out, 'rt') as f:
    assert len(f.read().splitlines()) == 3


def test_sweep_build_resumes_without_redrawing(tmp_path, monkeypatch):
  """``sweep.build`` records the chains it kept and skipped, and a resumed build draws only the
  indices it has not seen (each chain counted once)."""
  import json
  from tensorforge_amd.tune import sweep
  drawn = []

  def fake_job(args):
    i = args[0]
    drawn.append(i)
    rec = {'chain': i, 'info': {}, 'batch': 1000, 'dropped_large': [], 'candidates': []}
    if i % 3 == 0:
      rec['candidates'] = [{'label': 'plain'}, {'label': 'x'}]
      rec['so'] = f'libc{i}.so'
    return rec, 0.5, 1.0 if rec['candidates'] else 0.0

  monkeypatch.setattr(sweep, '_chain_job', fake_job)
  monkeypatch.setattr(sweep.concurrent.futures, 'ProcessPoolExecutor',
                      sweep.concurrent.futures.ThreadPoolExecutor)
  d = str(tmp_path / 'sw')
  sweep.build(d, 10, seed=1, jobs=2, start=100)
  sweep.build(d, 12, seed=1, jobs=2, start=100)
  assert sorted(drawn) == list(range(100, 112))
  idx = json.load(open(tmp_path / 'sw' / 'index.json'))
  kept = [c['chain'] for c in idx['chains']]
  assert kept == [102, 105, 108, 111]
  assert idx['skipped'] == len(idx['skipped_chains']) == 8
  assert idx['generate_seconds'] == 6.0 and idx['compile_seconds'] == 4.0
