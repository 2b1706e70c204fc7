"""Training the cost model on the sweep's timings, and judging its picks against the sweep.

The model predicts log(time) of a (chain, batch, candidate) from generation-time features
(``features.NAMES``). Used two ways:

* pick: the candidate with the smallest predicted time;
* pruned search: time only the ``k - 1`` best-predicted candidates and the library's default
  choice, keep the fastest measured one (the exhaustive sweep stays the fallback and the yardstick).

``evaluate`` reports, over chains held out of training, the slowdown of each method's choice
against the sweep's best (``time(choice) / time(best) - 1``) for the model's pick, the pruned
search, the library's default choice (``emitter='auto'``, default schedule) and the mean of a
random choice.
"""
import json
from collections import defaultdict

import numpy as np

DEFAULT_ORDER = ('stream_w8_k8_a2', 'cpl_w1_k16', 'plain')


def load(path):
  """Timing records of a database (JSON lines, gzip-compressed when the name ends in .gz)."""
  import gzip
  with (gzip.open(path, 'rt') if path.endswith('.gz') else open(path)) as f:
    return [json.loads(line) for line in f if line.strip()]


def by_chain(records):
  out = defaultdict(list)
  for r in records:
    out[r['chain']].append(r)
  return dict(out)


def split(records, test_fraction=0.25, seed=0):
  chains = sorted(by_chain(records))
  rng = np.random.default_rng(seed)
  rng.shuffle(chains)
  n_test = max(1, int(round(len(chains) * test_fraction)))
  test = set(chains[:n_test])
  return ([r for r in records if r['chain'] not in test],
          [r for r in records if r['chain'] in test])


def feature_matrix(records):
  """Generation-time features of each record, the indicators of its label's knobs and the
  occupancy and load-balance estimates derived from them."""
  from .features import derived_features, label_features
  return np.array([list(r['features']) + label_features(r['label']) +
                   derived_features(r['features']) for r in records], dtype=np.float64)


def arrays(records, centered=True, target='centered', cap=1.0):
  """Features and the training target from log(time).

  ``target='centered'`` (or ``centered=True``): log(time) minus the chain's mean: the choice
  among one chain's candidates only depends on their differences, so the network spends its
  capacity on what the schedule changes rather than on the chain's overall cost.
  ``target='gap'``: log(time / the chain's best time), capped at ``cap``: every candidate slower
  than e^cap times the best looks the same, so the network's accuracy goes to ranking the ones
  near the best (where a pick is decided). ``target='raw'``: log(time)."""
  X = feature_matrix(records)
  y = np.log(np.array([r['time'] for r in records], dtype=np.float64))
  if not centered:
    target = 'raw'
  if target in ('centered', 'gap'):
    chains = np.array([r['chain'] for r in records])
    for c in np.unique(chains):
      sel = chains == c
      y[sel] -= y[sel].mean() if target == 'centered' else y[sel].min()
    if target == 'gap':
      y = np.minimum(y, cap)
  return X, y


class Ensemble:
  """Several networks trained from different seeds; the prediction is their mean (lower
  variance than one network: the ranking of close candidates is what decides a pick)."""

  def __init__(self, models):
    self.models = list(models)

  def predict(self, X):
    return np.mean([m.predict(X) for m in self.models], axis=0)


def train(records, epochs=400, lr=2e-3, seed=0, device='cuda:0', centered=True, ensemble=1,
          batch=128, target='centered', cap=1.0):
  """A network (``ensemble`` = 1) or an :class:`Ensemble` of ``ensemble`` networks trained on
  ``records``; returns it and the training-loss history of the first network."""
  from .mlp import CostMLP
  X, y = arrays(records, centered, target, cap)
  models, history = [], None
  for k in range(ensemble):
    model = CostMLP(X.shape[1], seed=seed + 1000 * k, device=device)
    h = model.fit(X, y, epochs=epochs, lr=lr, batch=batch, seed=seed + 1000 * k)
    history = history if history is not None else h
    models.append(model)
  return (models[0] if ensemble == 1 else Ensemble(models)), history


def default_choice(labels):
  for name in DEFAULT_ORDER:
    if name in labels:
      return name
  return labels[0]


def evaluate(model, records, prune_k=2):
  """Slowdowns against the sweep's best on the chains of ``records``."""
  res = {'model_pick': [], 'pruned_search': [], 'default': [], 'random_mean': [],
         'candidates': [], 'timed_by_pruned_search': []}
  for chain, recs in sorted(by_chain(records).items()):
    times = np.array([r['time'] for r in recs])
    best = times.min()
    pred = model.predict(feature_matrix(recs))
    order = np.argsort(pred)
    labels = [r['label'] for r in recs]
    res['model_pick'].append(times[order[0]] / best - 1)
    k = min(prune_k, len(recs))
    res['pruned_search'].append(times[pruned(order, labels, k)].min() / best - 1)
    res['default'].append(times[labels.index(default_choice(labels))] / best - 1)
    res['random_mean'].append(float(np.mean(times / best - 1)))
    res['candidates'].append(len(recs))
    res['timed_by_pruned_search'].append(k)
  summary = {'chains': len(res['model_pick'])}
  for key in ('model_pick', 'pruned_search', 'default', 'random_mean'):
    v = np.array(res[key])
    summary[key] = {'mean_slowdown': float(v.mean()), 'median_slowdown': float(np.median(v)),
                    'within_5pct': float((v <= 0.05).mean()), 'max_slowdown': float(v.max())}
  summary['mean_candidates'] = float(np.mean(res['candidates']))
  summary['mean_timed_by_pruned_search'] = float(np.mean(res['timed_by_pruned_search']))
  return summary


def pruned(order, labels, k):
  """Indices the pruned search times: the model's k - 1 best plus the default choice."""
  picks = [int(i) for i in order[:max(1, k - 1)]]
  d = labels.index(default_choice(labels))
  if d not in picks:
    picks.append(d)
  for i in order:
    if len(picks) >= min(k, len(labels)):
      break
    if int(i) not in picks:
      picks.append(int(i))
  return picks


def cross_validate(records, folds=4, seed=0, prune_k=2, epochs=400, device='cuda:0',
                   centered=True, ensemble=1, batch=128, lr=2e-3, target='centered', cap=1.0):
  """Every chain held out once: ``folds`` models (each an ensemble of ``ensemble`` networks),
  each trained on the other chains' timings; the per-chain slowdowns of all folds are pooled
  into one summary."""
  chains = sorted(by_chain(records))
  rng = np.random.default_rng(seed)
  rng.shuffle(chains)
  parts = [set(chains[i::folds]) for i in range(folds)]

  pooled = []
  for i, test in enumerate(parts):
    train_recs = [r for r in records if r['chain'] not in test]
    test_recs = [r for r in records if r['chain'] in test]
    model, _ = train(train_recs, epochs=epochs, lr=lr, seed=seed + i, device=device,
                     centered=centered, ensemble=ensemble, batch=batch, target=target, cap=cap)
    pooled.append((model, test_recs))
  out = {'model_pick': [], 'pruned_search': [], 'default': [], 'random_mean': []}
  for model, recs in pooled:
    for chain, rs in sorted(by_chain(recs).items()):
      times = np.array([r['time'] for r in rs])
      best = times.min()
      order = np.argsort(model.predict(feature_matrix(rs)))
      labels = [r['label'] for r in rs]
      out['model_pick'].append(times[order[0]] / best - 1)
      out['pruned_search'].append(times[pruned(order, labels, prune_k)].min() / best - 1)
      out['default'].append(times[labels.index(default_choice(labels))] / best - 1)
      out['random_mean'].append(float(np.mean(times / best - 1)))
  summary = {'chains': len(out['model_pick']), 'folds': folds, 'prune_k': prune_k,
             'ensemble': ensemble, 'epochs': epochs, 'batch': batch, 'lr': lr,
             'target': target if centered else 'raw'}
  for key, v in out.items():
    v = np.array(v)
    summary[key] = {'mean_slowdown': float(v.mean()), 'median_slowdown': float(np.median(v)),
                    'within_5pct': float((v <= 0.05).mean()), 'max_slowdown': float(v.max())}
  return summary
