"""Autotuning the per-order chains of an uneven batch (orders 2..7, dense K_d).

    python -m tensorforge_amd.tune.orders build   --dir D      (CPU: hipcc)
    python -m tensorforge_amd.tune.orders measure --dir D --out F   (GPU)

Every order's chain is built with each candidate streaming schedule (8 wavefronts per workgroup,
so any choice can join one grouped launch) and timed at its element count in the uneven batch;
the fastest schedule per order is written to ``F`` (JSON). The catalog's grouped entry reads the
committed choice (``tune/data/mixed_orders_mi355x.json``).
"""
import argparse
import concurrent.futures
import itertools
import json
import os
import sys

from .. import workloads as W
from ..common.basic_types import FloatingPointType
from ..common.context import Context
from ..emit.tuned import Schedule
from ..generators.generator import Generator

CHOICE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data',
                           'mixed_orders_mi355x.json')


def candidates():
  out = [{}]
  for kb, la, pf in itertools.product((4, 8), (0, 2), (1, 2)):
    out.append({'stream_k_block': kb, 'stream_lds_lookahead': la, 'stream_prefetch': pf})
  for lkb, rows in itertools.product((2, 4, 8), (48, 64, 96)):
    out.append({'stream_loop_k_block': lkb, 'stream_pass_rows': rows})
  return out


def member(b, kw):
  w = W.ader_volume(FloatingPointType.DOUBLE, b=b, sparse_k=False)
  g = Generator(w.descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE), emitter='tuned',
                schedule=Schedule(**kw))
  g.generate()
  return w, g


def chosen_schedule(b):
  """The committed fastest schedule for basis size ``b`` (None: the default)."""
  if not os.path.exists(CHOICE_PATH):
    return None
  data = json.load(open(CHOICE_PATH))
  rec = data.get('orders', {}).get(str(b))
  return Schedule(**rec['schedule']) if rec else None


def build(directory, jobs=8):
  from ..runtime import build as B
  os.makedirs(directory, exist_ok=True)
  work, index = [], []
  for b in W.MIXED_BASES:
    seen = set()
    for kw in candidates():
      _, g = member(b, kw)
      em = g.get_emitted()
      if em.schedule.get('kind') != 'stream':
        continue
      key = g.get_source()
      if key in seen:          # knobs that do not apply to this order give the same kernel
        continue
      seen.add(key)
      name = f'b{b}_{len(index)}'
      index.append({'b': b, 'schedule': kw, 'so': f'lib{name}.so'})
      work.append((g, name))
  with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
    list(pool.map(lambda it: B.compile_generators([it[0]], it[1], directory), work))
  for f in os.listdir(directory):
    if f.endswith('.hip'):
      os.remove(os.path.join(directory, f))
  json.dump({'candidates': index}, open(os.path.join(directory, 'index.json'), 'w'))
  print(f'{len(index)} kernels', flush=True)


def measure(directory, out_path):
  import torch
  from ..runtime.loader import load_kernels
  from .measure import synthetic_buffers, time_launcher
  idx = json.load(open(os.path.join(directory, 'index.json')))
  counts = dict(zip(W.MIXED_BASES, W.mixed_counts()))
  dev = torch.device('cuda:0')
  res = {}
  for b in W.MIXED_BASES:
    n = counts[b]
    bufs = None
    ref = None
    cands = [c for c in idx['candidates'] if c['b'] == b]
    cands.sort(key=lambda c: c['schedule'] != {})          # the default schedule first
    for c in cands:
      _, g = member(b, c['schedule'])
      k = load_kernels(os.path.join(directory, c['so']), [g])[0]
      if bufs is None:
        bufs = synthetic_buffers(g, n, dev)
      try:
        bufs['Qp'].zero_()
        t = sorted(time_launcher(k, bufs, n, steps=10, warmup=2) for _ in range(3))[1]
      except RuntimeError as err:
        print(f'b={b} {c["schedule"]}: {err}', flush=True)
        continue
      # a candidate must reproduce the default schedule's result (the default is checked
      # against the oracle by the GPU tests); Qp is overwritten by every launch
      out = bufs['Qp'].double()
      if ref is None:
        ref = out.clone()
      diff = float(((out - ref).abs().amax() / ref.abs().amax().clamp_min(1e-300)).item())
      if diff > 1e-12:
        print(f'b={b} {c["schedule"]}: result differs from the default schedule ({diff:.2e})',
              flush=True)
        continue
      res.setdefault(str(b), []).append({'schedule': c['schedule'], 'time': t})
    del bufs
    torch.cuda.empty_cache()
    best = min(res[str(b)], key=lambda r: r['time'])
    default = next(r for r in res[str(b)] if r['schedule'] == {})
    print(f'b={b}: {len(res[str(b)])} schedules, best {best["time"] * 1e3:.3f} ms '
          f'{best["schedule"]}, default {default["time"] * 1e3:.3f} ms', flush=True)
  out = {'counts': {str(b): counts[b] for b in W.MIXED_BASES}, 'orders': {}, 'timings': res}
  for b, rs in res.items():
    best = min(rs, key=lambda r: r['time'])
    default = next(r for r in rs if r['schedule'] == {})
    out['orders'][b] = {'schedule': best['schedule'], 'time': best['time'],
                        'default_time': default['time']}
  json.dump(out, open(out_path, 'w'), indent=1)


def main(argv=None):
  ap = argparse.ArgumentParser()
  ap.add_argument('phase', choices=['build', 'measure'])
  ap.add_argument('--dir', required=True)
  ap.add_argument('--out', default=CHOICE_PATH)
  ap.add_argument('--jobs', type=int, default=8)
  a = ap.parse_args(argv)
  if a.phase == 'build':
    build(a.dir, a.jobs)
  else:
    measure(a.dir, a.out)


if __name__ == '__main__':
  sys.exit(main())
