"""Exhaustive sweep of the candidate schedules over random contraction chains: the timing
database the cost model is trained on, and the yardstick for its picks.

    python -m tensorforge_amd.tune.sweep build   --dir D [--chains N] [--seed S]   (CPU: hipcc)
    python -m tensorforge_amd.tune.sweep measure --dir D                           (GPU)

``build`` keeps the chains that admit at least two candidates, compiles every candidate into
its own shared object (the candidates of one chain share the launcher name) and writes
``D/index.json`` with each candidate's features. ``measure`` times every candidate on synthetic
inputs at the chain's batch size (HIP events, median of three runs of ``steps`` launches) and
appends one JSON line per candidate to ``D/timings.jsonl``.
"""
import argparse
import concurrent.futures
import json
import os
import sys
import time

from . import features as F
from . import space


def _chains(count, seed):
  return space.random_chains(count, seed)


def build(directory, count, seed, jobs=8, min_candidates=2):
  from ..runtime import build as B
  os.makedirs(directory, exist_ok=True)
  entries, work = [], []
  for i, (descrs, fp, batch, info) in enumerate(_chains(count, seed)):
    cands = space.applicable(descrs, fp)
    if len(cands) < min_candidates:
      continue
    rec = {'chain': i, 'info': info, 'batch': batch, 'candidates': []}
    for j, (c, g) in enumerate(cands):
      name = f'c{i}_{j}'
      rec['candidates'].append({'label': c.label, 'so': f'lib{name}.so',
                                'features': F.features(g, c, batch, info)})
      work.append((g, name))
    entries.append(rec)

  def one(item):
    g, name = item
    return B.compile_generators([g], name, directory)

  t0 = time.time()
  with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
    for k, _ in enumerate(pool.map(one, work)):
      if k % 25 == 0:
        print(f'compiled {k + 1}/{len(work)} ({time.time() - t0:.0f} s)', flush=True)
  for f in os.listdir(directory):
    if f.endswith('.hip'):
      os.remove(os.path.join(directory, f))
  with open(os.path.join(directory, 'index.json'), 'w') as f:
    json.dump({'count': count, 'seed': seed, 'feature_names': F.NAMES, 'chains': entries}, f)
  print(f'{len(entries)} chains, {len(work)} kernels', flush=True)


def measure(directory, steps=5, out_path=None):
  import torch
  from ..runtime.loader import load_kernels
  from .measure import synthetic_buffers, time_launcher
  idx = json.load(open(os.path.join(directory, 'index.json')))
  chains = _chains(idx['count'], idx['seed'])
  dev = torch.device('cuda:0')
  out_path = out_path or os.path.join(directory, 'timings.jsonl')
  with open(out_path, 'w') as out:
    for n, rec in enumerate(idx['chains']):
      descrs, fp, batch, info = chains[rec['chain']]
      cands = dict((c.label, (c, g)) for c, g in space.applicable(descrs, fp))
      bufs = None
      for cand in rec['candidates']:
        c, g = cands[cand['label']]
        k = load_kernels(os.path.join(directory, cand['so']), [g])[0]
        if bufs is None:
          bufs = synthetic_buffers(g, batch, dev)
        try:
          first = time_launcher(k, bufs, batch, steps=1, warmup=1)
        except RuntimeError as err:   # the launcher reported a launch error: not a candidate
          print(f'chain {rec["chain"]} {cand["label"]}: {err}', flush=True)
          continue
        n_steps = int(min(steps, max(1, 0.02 / max(first, 1e-7))))
        ts = sorted(time_launcher(k, bufs, batch, steps=n_steps, warmup=0) for _ in range(3))
        out.write(json.dumps({'chain': rec['chain'], 'label': cand['label'], 'time': ts[1],
                              'batch': batch, 'info': info, 'features': cand['features']}) + '\n')
      del bufs
      torch.cuda.empty_cache()
      print(f'chain {n + 1}/{len(idx["chains"])} ({info["kind"]}, {len(rec["candidates"])} '
            f'candidates) done', flush=True)


def main(argv=None):
  ap = argparse.ArgumentParser()
  ap.add_argument('phase', choices=['build', 'measure'])
  ap.add_argument('--dir', required=True)
  ap.add_argument('--chains', type=int, default=150)
  ap.add_argument('--seed', type=int, default=5)
  ap.add_argument('--jobs', type=int, default=8)
  ap.add_argument('--out', default=None, help='timings file (default: DIR/timings.jsonl)')
  a = ap.parse_args(argv)
  if a.phase == 'build':
    build(a.dir, a.chains, a.seed, a.jobs)
  else:
    measure(a.dir, out_path=a.out)


if __name__ == '__main__':
  sys.exit(main())
