"""The learned cost model against the exhaustive sweep (GPU: the network trains in HIP).

Data: ``tensorforge_amd/tune/data/timings_mi355x.jsonl``, every applicable candidate of 218
random contraction chains timed on an MI355X (``python -m tensorforge_amd.tune.sweep``, two
seeds). Five-fold cross-validation by chain: each chain is scored by a network that never saw it.

What the model (the mean of five networks trained from different seeds) gains over the method
without it (the library's default choice): its single pick is slower than the sweep's best by
less on average and lands within 5 % of it more often, and a search that times only the model's
two best candidates plus the default lands within 5 % of the sweep's best on most chains while
timing 3 of ~9.4 candidates.
"""
import os

import pytest

pytestmark = pytest.mark.gpu

DB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'tensorforge_amd',
                  'tune', 'data', 'timings_mi355x.jsonl')


def test_pruned_search_beats_the_default_choice(device):
  from tensorforge_amd.tune import model as M
  recs = M.load(DB)
  s = M.cross_validate(recs, folds=5, seed=0, prune_k=3, epochs=400, centered=True,
                       ensemble=5)
  print(s)
  cands = len(recs) / s['chains']
  assert s['chains'] == 218 and cands > 9
  assert s['model_pick']['mean_slowdown'] < s['default']['mean_slowdown']
  assert s['model_pick']['within_5pct'] >= s['default']['within_5pct'] + 0.05
  assert s['pruned_search']['within_5pct'] >= 0.85
  assert s['pruned_search']['within_5pct'] > s['default']['within_5pct'] + 0.1
  assert s['pruned_search']['mean_slowdown'] < 0.5 * s['default']['mean_slowdown']
  # any single pick is far better than a random one (plain kernels are several times slower)
  assert s['model_pick']['mean_slowdown'] < 0.2 * s['random_mean']['mean_slowdown']
