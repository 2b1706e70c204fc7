"""Exhaustive sweep of the candidate schedules over random contraction chains: the timing
database the cost model is trained on, and the yardstick for its picks.

    python -m tensorforge_amd.tune.sweep build   --dir D [--chains N] [--seed S]   (CPU: hipcc)
    python -m tensorforge_amd.tune.sweep measure --dir D                           (GPU)

``build`` draws ``N`` random chains and keeps the ones that admit at least two candidates. All
candidates of one chain go into one shared object under their own names
(``launcher_c<chain>_<label>``), so one hipcc call builds a chain's whole candidate set. A plain
or column-per-lane candidate whose generated source exceeds ``--max-source`` bytes is left out
(its unrolled C++ takes from half a minute to minutes to compile); the streaming kernels (inline
assembly, seconds to compile at any size) are always kept. Chains are
generated and compiled in worker processes; ``D/index.json`` records each chain's kept
candidates, features and launcher parameter list.

``measure`` needs no kernel generation: it rebuilds each chain's descriptions (only to name the
launcher parameters and size the synthetic inputs), loads the chain's object, and times every
candidate on synthetic inputs at the chain's batch size: one untimed launch, then the median of
per-launch HIP-event durations over up to ``steps`` launches (fewer for slow kernels). One JSON
line per candidate is appended to ``D/timings.jsonl``; a chain already present there is skipped,
so a run can be resumed.
"""
import argparse
import concurrent.futures
import json
import os
import sys
import time

from . import features as F
from . import space

MAX_SOURCE = 500_000


def _chain_job(args):
  """Worker: generate chain ``i``'s candidates and compile them into ``lib c<i>.so``."""
  i, seed_chain, directory, max_source, precision = args
  import numpy as np
  from ..runtime import build as B
  rng = np.random.default_rng([seed_chain, i])
  chain = None
  while chain is None:
    chain = space.random_chain(rng)
  descrs, fp, batch, info = chain
  t0 = time.time()
  # the tuned forms first: a chain none of them admits has nothing to choose (only the plain
  # kernel) and is skipped before the plain kernel is generated
  space_ = space.candidate_space(precision)
  tuned = space.applicable(descrs, fp, [c for c in space_ if c.emitter != 'plain'],
                           name_prefix=f'c{i}')
  cands = (space.applicable(descrs, fp, [c for c in space_ if c.emitter == 'plain'],
                            name_prefix=f'c{i}') + tuned) if tuned else []
  kept, dropped = [], []
  for c, g in cands:
    # streaming kernels are straight-line inline assembly and compile in seconds at any size;
    # the plain and column-per-lane kernels' unrolled C++ takes minutes above the cap
    big = not c.label.startswith('stream') and len(g.get_kernel()) > max_source
    (dropped if big else kept).append((c, g))
  rec = {'chain': i, 'info': info, 'batch': batch, 'dropped_large': [c.label for c, _ in dropped],
         'candidates': []}
  if len(kept) < 2:
    return rec, time.time() - t0, 0.0
  for c, g in kept:
    rec['candidates'].append({'label': c.label, 'launcher': f'launcher_{g.get_base_name()}',
                              'features': F.features(g, c, batch, info)})
  t1 = time.time()
  # device code stored compressed: thousands of objects travel to the GPU for timing
  B.compile_generators([g for _, g in kept], f'c{i}', directory,
                       extra_flags=['--offload-compress'])
  src = os.path.join(directory, f'c{i}.hip')
  if os.path.exists(src):
    os.remove(src)
  rec['so'] = f'libc{i}.so'
  return rec, t1 - t0, time.time() - t1


def chain_descriptions(seed_chain, i):
  """The descriptions of random chain ``i`` (same draw as the build)."""
  import numpy as np
  rng = np.random.default_rng([seed_chain, i])
  chain = None
  while chain is None:
    chain = space.random_chain(rng)
  return chain


def build(directory, count, seed, jobs=8, max_source=MAX_SOURCE, precision='exact', start=0):
  """Draw and compile chains ``start .. start + count - 1`` into ``directory``; resumable: the
  index records the chains kept and the chains skipped (fewer than two candidates), and a
  later call draws only the others."""
  os.makedirs(directory, exist_ok=True)
  index_path = os.path.join(directory, 'index.json')
  entries, skipped_set, gen_s, cc_s = [], set(), 0.0, 0.0
  if os.path.exists(index_path):
    old = json.load(open(index_path))
    # an index without the skipped chains' list (an older format) cannot tell which indices
    # were skipped: they are drawn again, and counted once
    entries, skipped_set = old['chains'], set(old.get('skipped_chains', []))
    gen_s, cc_s = old.get('generate_seconds', 0.0), old.get('compile_seconds', 0.0)
  done = {e['chain'] for e in entries} | skipped_set
  todo = [(i, seed, directory, max_source, precision) for i in range(start, start + count)
          if i not in done]
  t0 = time.time()

  def save():
    with open(index_path + '.tmp', 'w') as f:
      json.dump({'seed': seed, 'feature_names': F.NAMES, 'max_source': max_source,
                 'precision': precision, 'chains': sorted(entries, key=lambda e: e['chain']),
                 'skipped': len(skipped_set), 'skipped_chains': sorted(skipped_set),
                 'generate_seconds': gen_s, 'compile_seconds': cc_s}, f)
    os.replace(index_path + '.tmp', index_path)

  with concurrent.futures.ProcessPoolExecutor(jobs) as pool:
    for k, (rec, g_s, c_s) in enumerate(pool.map(_chain_job, todo, chunksize=1)):
      gen_s += g_s
      cc_s += c_s
      if rec['candidates']:
        entries.append(rec)
      else:
        skipped_set.add(rec['chain'])
      if k % 20 == 19 or k + 1 == len(todo):
        save()
        print(f'{k + 1}/{len(todo)} chains drawn, {len(entries)} kept '
              f'({time.time() - t0:.0f} s)', flush=True)
  save()


def _time_candidate(fn, args, steps, budget_s=0.02):
  """Median per-launch duration (s): one untimed launch, then up to ``steps`` timed ones (fewer
  when one launch takes more than ``budget_s / steps``)."""
  import torch
  stream = torch.cuda.current_stream()
  e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  e0.record(stream)
  fn(*args)
  e1.record(stream)
  torch.cuda.synchronize()
  first = e0.elapsed_time(e1) / 1e3
  n = int(min(steps, max(3, budget_s / max(first, 1e-7))))
  ev = [torch.cuda.Event(enable_timing=True) for _ in range(n + 1)]
  ev[0].record(stream)
  for j in range(n):
    fn(*args)
    ev[j + 1].record(stream)
  torch.cuda.synchronize()
  ts = sorted(ev[j].elapsed_time(ev[j + 1]) / 1e3 for j in range(n))
  return ts[len(ts) // 2] if n % 2 else 0.5 * (ts[n // 2 - 1] + ts[n // 2])


def measure(directory, steps=25, out_path=None, limit=None):
  import ctypes
  import torch
  from ..common.context import Context
  from ..device import launch_args
  from ..generators.generator import Generator
  from ..runtime.loader import Kernel, runtime
  from .measure import synthetic_buffers
  idx = json.load(open(os.path.join(directory, 'index.json')))
  dev = torch.device('cuda:0')
  out_path = out_path or os.path.join(directory, 'timings.jsonl')
  have = set()
  if os.path.exists(out_path):
    with open(out_path) as f:
      have = {json.loads(line)['chain'] for line in f if line.strip()}
  runtime()
  chains = [r for r in idx['chains'] if r['chain'] not in have]
  if limit is not None:
    chains = chains[:limit]
  t0 = time.time()
  with open(out_path, 'a') as out:
    for n, rec in enumerate(chains):
      descrs, fp, batch, info = chain_descriptions(idx['seed'], rec['chain'])
      g = Generator(descrs, Context('gfx950', 'hip', fp))
      g.register()
      lib = ctypes.CDLL(os.path.join(directory, rec['so']))
      bufs = synthetic_buffers(g, batch, dev)
      args = launch_args(g, {k: v for k, v in bufs.items() if k != '__keep__'}, batch)
      lines = []
      for cand in rec['candidates']:
        g.set_kernel_name(cand['launcher'][len('launcher_'):])
        k = Kernel(g, lib)
        try:
          t = _time_candidate(k, args, steps)
        except RuntimeError as err:   # the launcher reported a launch error: not a candidate
          print(f'chain {rec["chain"]} {cand["label"]}: {err}', flush=True)
          continue
        lines.append(json.dumps({'chain': rec['chain'], 'label': cand['label'], 'time': t,
                                 'batch': batch, 'info': info, 'features': cand['features']}))
      out.write(''.join(line + '\n' for line in lines))
      out.flush()
      del bufs, args
      torch.cuda.empty_cache()
      if n % 25 == 24 or n + 1 == len(chains):
        print(f'{n + 1}/{len(chains)} chains timed ({time.time() - t0:.0f} s)', flush=True)


def pack(directories, timing_paths, out_path, label):
  """One gzip-compressed timing database from sweep directories and their timing files:
  features rounded to 7 significant digits, times to 6, each record tagged with ``label``
  (``set``). A candidate timed in several files (separate measure passes) keeps the mean of its
  times and their count (``runs``): the sweep's best of a chain is then less often a lucky
  launch. Returns the sweep's throughput record: chains drawn, kept (at least two candidates)
  and timed, candidates timed, the build's CPU seconds (generation and compilation, summed over
  workers) from the directories' indexes."""
  import gzip
  times, first = {}, {}
  for path in timing_paths:
    with open(path) as f:
      for line in f:
        if line.strip():
          r = json.loads(line)
          key = (r['chain'], r['label'])
          times.setdefault(key, []).append(r['time'])
          first.setdefault(key, r)
  recs = []
  for key in sorted(first):
    r = dict(first[key])
    r['time'] = sum(times[key]) / len(times[key])
    r['runs'] = len(times[key])
    recs.append(r)
  with gzip.open(out_path, 'wt', compresslevel=9) as out:
    for r in recs:
      out.write(json.dumps({'chain': r['chain'], 'label': r['label'],
                            'time': float(f"{r['time']:.6g}"), 'batch': r['batch'],
                            'info': r['info'],
                            'features': [float(f'{v:.7g}') for v in r['features']],
                            'runs': r['runs'], 'set': label}, separators=(',', ':')) + '\n')
  meta = {'drawn': 0, 'kept': 0, 'generate_cpu_seconds': 0.0, 'compile_cpu_seconds': 0.0}
  for d in directories:
    idx = json.load(open(os.path.join(d, 'index.json')))
    meta['drawn'] += len(idx['chains']) + idx.get('skipped', 0)
    meta['kept'] += len(idx['chains'])
    meta['generate_cpu_seconds'] += idx.get('generate_seconds', 0.0)
    meta['compile_cpu_seconds'] += idx.get('compile_seconds', 0.0)
  meta['timed_chains'] = len({r['chain'] for r in recs})
  meta['timed_candidates'] = len(recs)
  return meta


def main(argv=None):
  ap = argparse.ArgumentParser()
  ap.add_argument('phase', choices=['build', 'measure', 'pack'])
  ap.add_argument('--dir', required=True, help='sweep directory (pack: comma-separated list)')
  ap.add_argument('--chains', type=int, default=150)
  ap.add_argument('--start', type=int, default=0, help='index of the first chain drawn')
  ap.add_argument('--seed', type=int, default=5)
  ap.add_argument('--jobs', type=int, default=8)
  ap.add_argument('--max-source', type=int, default=MAX_SOURCE)
  ap.add_argument('--precision', choices=['exact', 'mixed'], default='exact')
  ap.add_argument('--limit', type=int, default=None, help='time at most this many chains')
  ap.add_argument('--out', default=None, help='timings file (default: DIR/timings.jsonl; pack: '
                  'the database, .jsonl.gz)')
  ap.add_argument('--timings', default='', help='pack: comma-separated timing files')
  ap.add_argument('--logs', default='', help='pack: comma-separated measure logs (their last '
                  'line gives the pass\'s chains and GPU seconds)')
  ap.add_argument('--label', default='sweep', help='pack: the records\' set name')
  a = ap.parse_args(argv)
  if a.phase == 'build':
    build(a.dir, a.chains, a.seed, a.jobs, a.max_source, a.precision, a.start)
  elif a.phase == 'measure':
    measure(a.dir, out_path=a.out, limit=a.limit)
  else:
    meta = pack(a.dir.split(','), a.timings.split(','), a.out, a.label)
    passes = []
    for path in filter(None, a.logs.split(',')):
      with open(path) as f:
        last = [line for line in f if 'chains timed' in line][-1]
      done, secs = last.split('/')[0], last.split('(')[1].split()[0]
      passes.append({'chains': int(done), 'seconds': float(secs)})
    meta['measure_passes'] = passes
    meta['seed'] = a.seed
    with open(a.out.replace('.jsonl.gz', '.json'), 'w') as f:
      json.dump(meta, f, indent=1)
    print(json.dumps(meta))


if __name__ == '__main__':
  sys.exit(main())
