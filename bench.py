"""Benchmark: the fused ADER-DG volume chain on one MI355X.

Workload (the headline configuration): ``Q'_e = sum_d K_d Q_e S_{e,d}``, d = 1..3, order 6
(56 basis functions x 9 quantities), 262,144 elements, FP64; K_d 56x56 sparse (~40% random
pattern, CSR values), S_{e,d} 9x9 star matrices (24 nonzeros, packed). One step = one launch of
the generated launcher over the whole batch, inputs resident in HBM.

Prints one JSON line: elements/s of the tuned kernel, the plain (reference-mapping) kernel's
elements/s on the same inputs, the HBM roofline of the tuned kernel (algorithmic bytes per launch
over its average duration, from HIP events on the launch stream), the FP32 variant, and the CPU
baseline (the OpenMP restatement of the reference's generated kernel, on a bounded sample, with
the GPU result's error against it).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  # the clocks settle over the first few dozen launches: 50 untimed launches reach the steady
  # state (5 read about 6 % slower on the headline chain)
  ap.add_argument('--steps', type=int, default=200)
  ap.add_argument('--warmup', type=int, default=50)
  ap.add_argument('--elements', type=int, default=262144)
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--cpu-seconds', type=float, default=25.0)
  ap.add_argument('--cpu-full', action='store_true',
                  help='time the CPU baseline over the whole batch (minutes) instead of a '
                       'bounded sample')
  ap.add_argument('--only', default='',
                  help='comma list of legs (tuned,plain,f32,copy,mixed,precision,flux,deriv,'
                       'tuner,pointwise,gemm)')
  return ap.parse_args()


def make_inputs(entry, n, device, dtype, seed=2):
  """Seeded device inputs in the launcher layout (values as in tensorforge_amd.workloads)."""
  import torch
  from tensorforge_amd import workloads as W
  g = torch.Generator(device=device)
  g.manual_seed(seed)
  b, q = entry.meta['b'], entry.meta['q']
  bufs = {}
  for d in range(3):
    K = entry.tensors[f'K{d + 1}']
    nnz = W.packed_count(K)
    bufs[f'K{d + 1}'] = (torch.rand(nnz, generator=g, device=device, dtype=torch.float64) * 2 - 1).to(dtype)
  bufs['Q'] = torch.randn(n * b * q, generator=g, device=device, dtype=torch.float64).to(dtype)
  nnz_s = W.packed_count(entry.tensors['S1'])
  for d in range(3):
    s = torch.zeros(n, q * q, device=device, dtype=dtype)
    s[:, :nnz_s] = (torch.rand(n, nnz_s, generator=g, device=device, dtype=torch.float64) * 4 - 2).to(dtype)
    bufs[f'S{d + 1}'] = s.reshape(-1)
  bufs['Qp'] = torch.zeros(n * b * q, device=device, dtype=dtype)
  return bufs


def step_times(fn, steps, warmup, stream):
  """Durations (s) of ``steps`` calls of ``fn`` after ``warmup`` untimed ones: one HIP event on
  ``stream`` before the first timed call and one after each, so consecutive events bracket one
  step each."""
  import torch
  for _ in range(warmup):
    fn()
  torch.cuda.synchronize()
  events = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
  events[0].record(stream)
  for i in range(steps):
    fn()
    events[i + 1].record(stream)
  torch.cuda.synchronize()
  return [events[i].elapsed_time(events[i + 1]) / 1e3 for i in range(steps)]


def median(xs):
  ys = sorted(xs)
  m = len(ys) // 2
  return ys[m] if len(ys) % 2 else 0.5 * (ys[m - 1] + ys[m])


def launch_times(kernel, bufs, n, steps, warmup):
  """Per-launch durations (s) of the launcher on the current stream."""
  import torch
  from tensorforge_amd.device import launch_args
  stream = torch.cuda.current_stream()
  args = launch_args(kernel.generator, bufs, n, stream=stream)
  return step_times(lambda: kernel(*args), steps, warmup, stream)


def time_kernel(kernel, bufs, n, steps, warmup):
  """Median per-launch duration (s) over ``steps`` launches after ``warmup``."""
  return median(launch_times(kernel, bufs, n, steps, warmup))


def copy_bandwidth(device, steps, warmup):
  """Measured HBM roof: 2 * bytes / median launch time of a 16-byte-vector copy of 2 GiB."""
  import torch
  from tensorforge_amd.runtime import runtime
  rt = runtime()
  n = (1 << 31) // 8
  src = torch.empty(n, dtype=torch.float64, device=device).normal_()
  dst = torch.empty_like(src)
  stream = torch.cuda.current_stream()
  t = median(step_times(
    lambda: rt.tensorforge_copy(dst.data_ptr(), src.data_ptr(), n, stream.cuda_stream),
    steps, warmup, stream))
  del src, dst
  return 2 * n * 8 / t / 1e9


def mixed_inputs(group, counts, device, seed=3):
  """Per-order device buffers of the uneven batch: Q ~ N(0,1), star values U[-2,2], dense
  K_d U[-1,1], Q' zero."""
  import torch
  from tensorforge_amd import workloads as W
  g = torch.Generator(device=device)
  g.manual_seed(seed)
  out = []
  for e, n in zip(group.members, counts):
    b, q = e.meta['b'], e.meta['q']
    bufs = {f'K{d + 1}': torch.rand(b * b, generator=g, device=device, dtype=torch.float64) * 2 - 1
            for d in range(3)}
    bufs['Q'] = torch.randn(n * b * q, generator=g, device=device, dtype=torch.float64)
    nnz_s = W.packed_count(e.tensors['S1'])
    for d in range(3):
      sv = torch.zeros(n, q * q, device=device, dtype=torch.float64)
      sv[:, :nnz_s] = torch.rand(n, nnz_s, generator=g, device=device, dtype=torch.float64) * 4 - 2
      bufs[f'S{d + 1}'] = sv.reshape(-1)
    bufs['Qp'] = torch.zeros(n * b * q, device=device, dtype=torch.float64)
    out.append(bufs)
  return out


def mixed_leg(device, steps, warmup, graph_steps=10):
  """Uneven batch (orders 2..7, 500,000 elements, dense K_d, FP64): one step as plain
  per-order launches, tuned per-order launches (each order with its autotuned schedule,
  tensorforge_amd/tune/data/mixed_orders_mi355x.json) and one grouped persistent launch; and
  ``graph_steps`` steps with Q and Q' swapping roles captured as one HIP graph (grouped launches,
  and tuned per-order launches)."""
  import torch
  from tensorforge_amd import catalog
  from tensorforge_amd import workloads as W
  from tensorforge_amd.device import launch_args
  from tensorforge_amd.graph import Graph
  group, tuned, plain, gk = catalog.load_group('ader_mixed_f64')
  counts = W.mixed_counts()
  total = sum(counts)
  bufs = mixed_inputs(group, counts, device)
  stream = torch.cuda.Stream()

  def member_args(swap):
    out = []
    for e, b, n in zip(group.members, bufs, counts):
      bb = dict(b, Q=b['Qp'], Qp=b['Q']) if swap else b
      out.append(launch_args(e.generator, bb, n, stream=stream))
    return out

  args = [member_args(False), member_args(True)]
  gargs = [group.grouped.member_args(a) for a in args]

  def timed(fn, n_steps, n_warm):
    with torch.cuda.stream(stream):
      return median(step_times(fn, n_steps, n_warm, stream))

  res = {'elements': total, 'counts': dict(zip((str(b) for b in W.MIXED_BASES), counts))}
  t_plain = timed(lambda: [k(*a) for k, a in zip(plain, args[0])], steps, warmup)
  t_tuned = timed(lambda: [k(*a) for k, a in zip(tuned, args[0])], steps, warmup)
  t_group = timed(lambda: gk(*gargs[0]), steps, warmup)

  def seq_grouped():
    for s in range(graph_steps):
      gk(*gargs[s % 2])

  def seq_members():
    for s in range(graph_steps):
      for k, a in zip(tuned, args[s % 2]):
        k(*a)

  graphs = {}
  for name, fn in (('grouped', seq_grouped), ('per_order', seq_members)):
    gr = Graph(stream)
    gr.capture(fn)
    t = timed(gr.launch, max(2, steps // graph_steps), max(1, warmup // graph_steps)) / graph_steps
    graphs[name] = t
    gr.close()
  res.update({
    'plain_per_order': {'value': total / t_plain, 'ms_per_step': t_plain * 1e3},
    'tuned_per_order': {'value': total / t_tuned, 'ms_per_step': t_tuned * 1e3},
    'grouped': {'value': total / t_group, 'ms_per_step': t_group * 1e3},
    'graph_grouped': {'value': total / graphs['grouped'], 'ms_per_step': graphs['grouped'] * 1e3,
                      'steps_per_graph': graph_steps},
    'graph_tuned_per_order': {'value': total / graphs['per_order'],
                              'ms_per_step': graphs['per_order'] * 1e3,
                              'steps_per_graph': graph_steps},
    'grouped_over_plain': t_plain / t_group,
    'unit': 'elements/s'})
  del bufs
  return res


def _normwise(got, ref, vol):
  import numpy as np
  g = got.reshape(-1, vol)
  r = ref.reshape(-1, vol)
  den = np.maximum(np.abs(r).max(axis=1), np.finfo(np.float64).tiny)
  return float((np.abs(g - r).max(axis=1) / den).max())


def precision_leg(device, steps, warmup, n=262144):
  """Order-7 chain (84 basis functions, dense K_d), |Q| log-uniform in [1e-6, 1e3]: FP64
  vector (streaming schedule), FP32 MFMA and split BF16x3 MFMA. Error: normwise per element
  against the FP64 kernel's result on all elements (the FP64 kernel itself is checked against
  the FP64 oracle to 1e-12 by the GPU tests)."""
  import numpy as np
  import torch
  from tensorforge_amd import catalog
  from tensorforge_amd import workloads as W
  b, q = 84, 9
  rng = np.random.default_rng(4)
  host = {'Q': W.log_uniform_q(rng, n * b * q)}
  for d in range(3):
    host[f'K{d + 1}'] = rng.uniform(-1, 1, b * b)
    sv = np.zeros((n, q * q))
    sv[:, :24] = rng.uniform(-2, 2, (n, 24))
    host[f'S{d + 1}'] = sv.reshape(-1)
  bufs = {k: torch.from_numpy(v).to(device) for k, v in host.items()}
  vol = b * q
  res = {'workload': 'order 7 (84x9), dense K_d, 262144 elements, |Q| log-uniform [1e-6, 1e3]',
         'error': 'normwise per element vs the FP64 kernel, all elements'}
  out64 = None
  for name, label in (('ader7_f64_tuned', 'fp64_vector'), ('ader7_f64_mfma_fp32', 'fp32_mfma'),
                      ('ader7_f64_mfma_bf16x3', 'bf16x3_mfma')):
    _, k = catalog.load(name)
    bufs['Qp'] = torch.zeros(n * vol, dtype=torch.float64, device=device)
    t = time_kernel(k, bufs, n, steps, warmup)
    got = bufs['Qp'].cpu().numpy()
    if out64 is None:
      out64 = got
    res[label] = {'value': n / t, 'unit': 'elements/s', 'ms_per_step': t * 1e3,
                  'err_vs_fp64_kernel': _normwise(got, out64, vol)}
  del bufs
  return res


def flux_leg(device, steps, warmup, n=262144):
  """Local and neighbour flux integrals (reference benchmarks/fused/programs/s3-*-integral.cf,
  operand order 'mod'), FP64, tuned and plain launchers on the same inputs."""
  import torch
  from tensorforge_amd import catalog
  from tensorforge_amd.tune.measure import synthetic_buffers
  res = {}
  for kind in ('local', 'neigh'):
    entry, tuned = catalog.load(f'flux_{kind}_mod_f64_tuned')
    _, plain = catalog.load(f'flux_{kind}_mod_f64_plain')
    bufs = synthetic_buffers(tuned.generator, n, device)
    bufs.pop('__keep__')
    bpe = tuned.generator.get_emitted().bytes_per_element
    t_t = time_kernel(tuned, bufs, n, steps, warmup)
    t_p = time_kernel(plain, bufs, n, steps, warmup)
    res[kind] = {'tuned': {'value': n / t_t, 'ms_per_step': t_t * 1e3, 'hbm_GBps': bpe * n / t_t / 1e9},
                 'plain': {'value': n / t_p, 'ms_per_step': t_p * 1e3},
                 'tuned_over_plain': t_p / t_t, 'bytes_per_element': bpe}
    del bufs
  res['unit'] = 'elements/s'
  res['elements'] = n
  return res


def deriv_leg(device, steps, warmup, n=262144):
  """Time-derivative step (reference benchmarks/fused/programs/seissol-deriv_align.cf and
  seissol-deriv_no_align.cf), FP64, on the same inputs: the default emitter choice (streaming
  when aligned; plain when not aligned, where the column-per-lane form holds one wavefront per
  CU) against the plain launcher, with the HBM roofline of the algorithmic bytes."""
  from tensorforge_amd import catalog
  from tensorforge_amd import workloads as W
  from tensorforge_amd.common.context import Context
  from tensorforge_amd.generators.generator import Generator
  from tensorforge_amd.tune.measure import synthetic_buffers
  res = {}
  for label, stem in (('aligned', 'deriv'), ('not_aligned', 'deriv_na')):
    entry, tuned = catalog.load(f'{stem}_f64_tuned')
    _, plain = catalog.load(f'{stem}_f64_plain')
    fresh = W.ader_derivative(entry.fp_type, aligned=(label == 'aligned'))
    default = Generator(fresh.descrs, Context('gfx950', 'hip', entry.fp_type))
    default.generate()
    choice = default.get_emitted().emitter
    bufs = synthetic_buffers(tuned.generator, n, device)
    bufs.pop('__keep__')
    em = tuned.generator.get_emitted()
    bpe = em.bytes_per_element
    # the column-per-lane form of the not-aligned chain takes ~30 ms per launch: fewer launches
    stream_form = em.schedule.get('kind') == 'stream'
    t_t = time_kernel(tuned, bufs, n, steps if stream_form else 3, warmup if stream_form else 2)
    t_p = time_kernel(plain, bufs, n, steps, warmup)
    t_d = t_t if choice == 'tuned' else t_p
    form = 'streaming' if em.schedule.get('kind') == 'stream' else 'column-per-lane'
    res[label] = {'default': {'emitter': choice if choice == 'plain' else form,
                              'value': n / t_d, 'ms_per_step': t_d * 1e3,
                              'hbm_GBps': bpe * n / t_d / 1e9,
                              'frac_of_peak': bpe * n / t_d / 1e9 / HBM_PEAK_GBPS},
                  'tuned': {'form': form, 'ms_per_step': t_t * 1e3},
                  'plain': {'value': n / t_p, 'ms_per_step': t_p * 1e3},
                  'default_over_plain': t_p / t_d, 'bytes_per_element': bpe}
    del bufs
  res['unit'] = 'elements/s'
  res['elements'] = n
  return res


def gemm_leg(device, steps, warmup, cpu_seconds, with_cpu, n=65536):
  """Single batched GEMM C_e += A * B_e (A 56x56 shared, B_e and C_e 56x9, seed 0, FP64,
  65,536 elements): tuned (streaming), plain and FP64 matrix-core launchers with the HBM
  roofline (12,096 B per element), each result against the closed form on every element, and the OpenMP restatement of the
  reference's generated kernel (oracle/omp_chain.cpp, team and flat forms) on the host."""
  import numpy as np
  import torch
  from tensorforge_amd import catalog
  from tensorforge_amd import workloads as W
  from tensorforge_amd.common.basic_types import FloatingPointType
  entry, tuned = catalog.load('gemm_f64_tuned')
  _, plain = catalog.load('gemm_f64_plain')
  _, mfma64 = catalog.load('gemm_f64_mfma_fp64')
  w = W.gemm(FloatingPointType.DOUBLE)
  host = W.gemm_inputs(w, n, seed=0)
  b, q = w.meta['b'], w.meta['q']
  want = host['C'].reshape(n, q, b) + np.einsum('ik,ekj->eji', host['A'].reshape(b, b).T,
                                                host['B'].reshape(n, q, b).transpose(0, 2, 1))
  res = {'elements': n, 'unit': 'elements/s'}
  em = tuned.generator.get_emitted()
  bpl = em.bytes_per_element * n + em.shared_bytes
  for label, k in (('tuned', tuned), ('plain', plain), ('mfma_fp64', mfma64)):
    bufs = {key: torch.from_numpy(v.copy()).to(device) for key, v in host.items()}
    from tensorforge_amd.device import launch
    launch(k, bufs, n)
    torch.cuda.synchronize()
    got = bufs['C'].cpu().numpy().reshape(n, q, b)
    err = float((np.abs(got - want).reshape(n, -1).max(axis=1) /
                 np.abs(want).reshape(n, -1).max(axis=1)).max())
    t = time_kernel(k, bufs, n, steps, warmup)
    res[label] = {'value': n / t, 'ms_per_step': t * 1e3, 'hbm_GBps': bpl / t / 1e9,
                  'frac_of_peak': bpl / t / 1e9 / HBM_PEAK_GBPS,
                  'max_normwise_rel_err_vs_closed_form': err}
    del bufs
  res['tuned_over_plain'] = res['plain']['ms_per_step'] / res['tuned']['ms_per_step']
  res['mfma_fp64_over_plain'] = res['plain']['ms_per_step'] / res['mfma_fp64']['ms_per_step']
  res['schedule'] = em.schedule
  res['mfma_fp64_schedule'] = mfma64.generator.get_emitted().schedule
  if with_cpu:
    lib = ctypes.CDLL(os.path.join(ROOT, 'oracle', 'lib', 'libomp_chain.so'))
    DP = ctypes.POINTER(ctypes.c_double)
    cpu = {'cores': lib.omp_max_threads(), 'kind': 'port',
           'host': host_info(lib.omp_max_threads())}
    for form in ('team', 'flat'):
      m = 256
      while True:
        C = host['C'][:m * b * q].copy()
        t0 = time.perf_counter()
        if form == 'team':
          lib.omp_gemm_team(host['A'].ctypes.data_as(DP), host['B'].ctypes.data_as(DP),
                            C.ctypes.data_as(DP), b, q, ctypes.c_long(m), 32, 8)
        else:
          lib.omp_gemm_flat(host['A'].ctypes.data_as(DP), host['B'].ctypes.data_as(DP),
                            C.ctypes.data_as(DP), b, q, ctypes.c_long(m))
        t = time.perf_counter() - t0
        if t > cpu_seconds / 4 or m == n:
          break
        m = min(n, m * 4)
      ref = C.reshape(m, q, b)
      cpu[form] = {'value': m / t, 'elements': m, 'seconds': t,
                   'max_normwise_rel_err_vs_closed_form': float(
                     (np.abs(ref - want[:m]).reshape(m, -1).max(axis=1) /
                      np.abs(want[:m]).reshape(m, -1).max(axis=1)).max())}
    # this package's omptarget launcher of the same GEMM (Context(backend='omptarget')), over
    # the whole batch
    _, gen, omp_fn = catalog.load_omp('gemm_f64')
    C = host['C'].copy()
    args, _keep = catalog.omp_arguments(gen, {'A': host['A'], 'B': host['B'], 'C': C}, n)
    t0 = time.perf_counter()
    omp_fn(*args)
    t = time.perf_counter() - t0
    cpu['generated_omptarget'] = {
      'value': n / t, 'elements': n, 'seconds': t,
      'max_normwise_rel_err_vs_closed_form': float(
        (np.abs(C.reshape(n, q, b) - want).reshape(n, -1).max(axis=1) /
         np.abs(want).reshape(n, -1).max(axis=1)).max())}
    res['cpu_baseline'] = cpu
  return res


def pointwise_leg(device, steps, warmup, n=262144):
  """Pointwise kernels (ElementwiseDescr) at the headline element count: the update
  Q += dt * I (FP64 and FP32, 2 reads + 1 write per entry) and the plastic return map on the six
  stress columns (FP64; after the first launch every node is on or inside the yield surface, so
  the timed launches store only the nodes that round outside it and ``hbm_GBps`` counts its
  loads only). For the update ``hbm_GBps`` counts every entry's two loads and one store."""
  from tensorforge_amd import catalog
  from tensorforge_amd.tune.measure import synthetic_buffers
  res = {}
  for name in ('pw_update_f64', 'pw_update_f32', 'pw_plastic_f64'):
    _, k = catalog.load(name)
    bufs = synthetic_buffers(k.generator, n, device)
    bufs.pop('__keep__')
    if name.startswith('pw_plastic'):
      bufs['cohesion'], bufs['tanphi'] = 1.2, 0.2
    em = k.generator.get_emitted()
    t = time_kernel(k, bufs, n, steps, warmup)
    # the return map is idempotent: once every node is on or inside the surface, a launch
    # stores (almost) nothing, so only its loads count
    bpe = em.schedule['load_bytes_per_element'] if name.startswith('pw_plastic') else \
        em.bytes_per_element
    res[name] = {'value': n / t, 'ms_per_step': t * 1e3, 'bytes_per_element': bpe,
                 'hbm_GBps': bpe * n / t / 1e9, 'frac_of_peak': bpe * n / t / 1e9 / HBM_PEAK_GBPS}
    del bufs
  # fusion: C = A B followed by C = max(C, 0) * dt as one kernel, against the two kernels
  _, fused = catalog.load('t_fused_gemm_pw')
  _, gemm = catalog.load('gemm_f64_tuned')
  _, clamp = catalog.load('pw_clamp_scale_f64')
  bufs = synthetic_buffers(fused.generator, n, device)
  bufs.pop('__keep__')
  t_fused = time_kernel(fused, bufs, n, steps, warmup)
  t_gemm = time_kernel(gemm, bufs, n, steps, warmup)
  t_clamp = time_kernel(clamp, bufs, n, steps, warmup)
  res['fused_gemm_statements'] = {'fused_ms': t_fused * 1e3, 'gemm_ms': t_gemm * 1e3,
                                  'statements_ms': t_clamp * 1e3,
                                  'split_over_fused': (t_gemm + t_clamp) / t_fused}
  del bufs
  res['unit'] = 'elements/s'
  res['elements'] = n
  return res


def tuner_leg():
  """Cost model vs the exhaustive sweep: five-fold cross-validation by chain over the
  committed timing database (20,000 random chains drawn; every applicable candidate of the 7,921
  that admit a choice timed twice on MI355X); the networks train on the GPU in the project's HIP
  kernels. Also the sweep's record: chains drawn and kept, the build's CPU seconds and the
  measure passes' GPU seconds."""
  from tensorforge_amd.tune import model as M
  from tensorforge_amd.tune.autotune import SETTINGS
  data = os.path.join(ROOT, 'tensorforge_amd', 'tune', 'data')
  recs = M.load(os.path.join(data, 'timings_mi355x.jsonl.gz'))
  t0 = time.time()
  s = M.cross_validate(recs, folds=5, seed=0, prune_k=3, **SETTINGS)
  s['cross_validation_seconds'] = time.time() - t0
  s['candidates_per_chain'] = len(recs) / s['chains']
  with open(os.path.join(data, 'timings_mi355x.json')) as f:
    s['sweep'] = json.load(f)
  return s


def profiled_traffic(kernel_name, schedule):
  """HBM bytes per launch of ``kernel_name`` from the committed rocprofv3 counter summary
  (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor the same counter reads for
  the copy kernel), when that summary was taken with the same schedule; else None."""
  path = os.path.join(ROOT, 'profiles', f'{kernel_name}_hbm_traffic.md')
  if not os.path.exists(path):
    return None
  fields = {}
  with open(path, encoding='utf-8') as f:
    for line in f:
      if ':' in line and line.startswith('- '):
        key, _, val = line[2:].partition(':')
        fields[key.strip()] = val.strip().strip('`')
  if fields.get('schedule') != json.dumps(schedule, sort_keys=True) or \
      'traffic_bytes_per_launch' not in fields:
    return None
  return int(fields['traffic_bytes_per_launch'])


def host_info(threads):
  """The host the CPU baseline ran on: CPU model, logical CPUs, CPUs this process may use, the
  OpenMP thread count and every OMP_* setting of the environment."""
  model = None
  try:
    with open('/proc/cpuinfo', encoding='utf-8') as f:
      for line in f:
        if line.startswith('model name'):
          model = line.split(':', 1)[1].strip()
          break
  except OSError:
    pass
  return {'cpu_model': model, 'logical_cpus': os.cpu_count(),
          'cpus_usable': len(os.sched_getaffinity(0)), 'omp_threads': threads,
          'omp_env': {k: v for k, v in sorted(os.environ.items()) if k.startswith('OMP_')}}


def cpu_baseline(entry, bufs, gpu_out, budget_s, full=False):
  """The OpenMP restatement of the reference's generated kernel (oracle/omp_chain.cpp, team
  form and flat form) on the first elements of the same inputs, sized to ~budget_s of CPU work;
  also the GPU result's max normwise error against it on that sample, and the same chain's
  launcher from this package's omptarget backend (``generated``: what ``Context(backend=
  'omptarget')`` builds from the same descriptions, bit-identical to the flat form)."""
  import numpy as np
  from tensorforge_amd import workloads as W
  lib = ctypes.CDLL(os.path.join(ROOT, 'oracle', 'lib', 'libomp_chain.so'))
  threads = lib.omp_max_threads()
  b, q = entry.meta['b'], entry.meta['q']
  host = {k: v.cpu().numpy().astype(np.float64) for k, v in bufs.items() if k != 'Qp'}
  masks = entry.meta['masks']
  rowptr, cols, vals = [], [], []
  for d in range(3):
    ents = W.csr_entries(masks[d])
    rp = np.zeros(b + 1, dtype=np.int32)
    for r, _ in ents:
      rp[r + 1] += 1
    rowptr.append(np.cumsum(rp).astype(np.int32))
    cols.append(np.array([c for _, c in ents], dtype=np.int32))
    vals.append(np.ascontiguousarray(host[f'K{d + 1}']))
  star = W.star_entries()
  srow = np.array([r for r, _ in star], dtype=np.int32)
  scol = np.array([c for _, c in star], dtype=np.int32)
  IP = ctypes.POINTER(ctypes.c_int)
  DP = ctypes.POINTER(ctypes.c_double)
  arr_i = (IP * 3)(*[x.ctypes.data_as(IP) for x in rowptr])
  arr_c = (IP * 3)(*[x.ctypes.data_as(IP) for x in cols])
  arr_v = (DP * 3)(*[x.ctypes.data_as(DP) for x in vals])
  S = [np.ascontiguousarray(host[f'S{d + 1}']) for d in range(3)]
  arr_s = (DP * 3)(*[x.ctypes.data_as(DP) for x in S])
  Q = np.ascontiguousarray(host['Q'])

  from tensorforge_amd import catalog
  _, gen, omp_fn = catalog.load_omp('ader_f64')

  def run(form, n):
    out = np.zeros(n * b * q)
    t0 = time.perf_counter()
    if form == 'generated':
      args, _keep = catalog.omp_arguments(gen, dict(host, Qp=out), n)
      t0 = time.perf_counter()
      omp_fn(*args)
    elif form == 'team':
      lib.omp_ader_team(arr_i, arr_c, arr_v, Q.ctypes.data_as(DP), arr_s, out.ctypes.data_as(DP),
                        srow.ctypes.data_as(IP), scol.ctypes.data_as(IP), len(star), b, q,
                        ctypes.c_long(n), 32, 8)
    else:
      lib.omp_ader_flat(arr_i, arr_c, arr_v, Q.ctypes.data_as(DP), arr_s, out.ctypes.data_as(DP),
                        srow.ctypes.data_as(IP), scol.ctypes.data_as(IP), len(star), b, q,
                        ctypes.c_long(n))
    return time.perf_counter() - t0, out

  total = len(Q) // (b * q)
  result = {'cores': threads, 'kind': 'port', 'host': host_info(threads)}
  flat_out = None
  for form in ('team', 'flat', 'generated'):
    if full:
      n = total
    else:
      n = 256
      t, _ = run(form, n)
      while t < budget_s / 8 and n < total:
        n = min(total, n * 4)
        t, _ = run(form, n)
      n = int(min(total, max(256, n * budget_s / 2 / max(t, 1e-6))) // 8 * 8)
    t, out = run(form, n)
    result[form] = {'value': n / t, 'elements': n, 'seconds': t}
    if form == 'generated':
      result[form]['bitwise_equal_to_flat'] = bool(
          np.array_equal(out[:min(n, flat_out[1]) * b * q], flat_out[0][:min(n, flat_out[1]) * b * q]))
    if form == 'flat':
      flat_out = (out, n)
      ref = out.reshape(n, b * q)
      got = gpu_out[:n * b * q].reshape(n, b * q)
      err = np.abs(got - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-300)
      result['gpu_max_normwise_rel_err'] = float(err.max())
  return result


def main():
  args = parse()
  if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
    print('bench.py measures one GPU: the launchers have no multi-GPU path (elements are independent; '
          'a batch splits into contiguous per-GPU ranges). Run with --gpus 1.', file=sys.stderr)
    sys.exit(2)
  import numpy as np
  import torch
  from tensorforge_amd import catalog
  device = torch.device('cuda:0')
  legs = set(args.only.split(',')) if args.only else {'tuned', 'plain', 'f32', 'copy', 'mixed',
                                                      'precision', 'flux', 'deriv', 'tuner',
                                                      'pointwise',
                                                      'gemm'}
  n = args.elements
  out = {}

  entry, tuned = catalog.load('ader_f64_tuned')
  bufs = make_inputs(entry, n, device, torch.float64)
  emitted = tuned.generator.get_emitted()
  bytes_per_launch = emitted.bytes_per_element * n + emitted.shared_bytes
  flops = tuned.generator.get_program().flops_per_element() * n

  # the headline kernel is timed after the other FP64-heavy legs (mixed orders, precision forms,
  # flux integrals, derivative step), as it runs among them in a time step: the GPU's power
  # management has then settled. Started from a lighter load, its first launches run at
  # 0.67 ms, the next ones are throttled to about 0.94 ms, and it takes a few dozen launches to
  # settle near 0.70-0.73 ms.
  measured_bw = copy_bandwidth(device, args.steps, args.warmup) if 'copy' in legs else None

  if 'plain' in legs:
    _, plain = catalog.load('ader_f64_plain')
    bufs_plain = dict(bufs)
    bufs_plain['Qp'] = torch.zeros_like(bufs['Qp'])
    t_plain = time_kernel(plain, bufs_plain, n, args.steps, args.warmup)
    plain_out = bufs_plain['Qp'].cpu().numpy().reshape(n, -1)
    del bufs_plain
    _, plain_dense = catalog.load('ader_dense_f64_plain')
    bufs_pd = dict(bufs)
    # the dense-K plain kernel reads K column-major: same values, densified
    from tensorforge_amd import workloads as W
    for d in range(3):
      dense = torch.zeros(56 * 56, dtype=torch.float64, device=device)
      ents = W.csr_entries(entry.meta['masks'][d])
      idx = torch.tensor([r + 56 * c for r, c in ents], device=device)
      dense[idx] = bufs[f'K{d + 1}']
      bufs_pd[f'K{d + 1}'] = dense
    bufs_pd['Qp'] = torch.zeros_like(bufs['Qp'])
    t_pd = time_kernel(plain_dense, bufs_pd, n, args.steps, args.warmup)
    pd_out = bufs_pd['Qp'].cpu().numpy().reshape(n, -1)
    # the same dense-K chain in the streaming form (one DPP FMA per entry, zeros included) and in
    # the FP64 matrix-core form, each against the plain dense-K result
    dense_forms = {}
    for label, name in (('tuned', 'ader_dense_f64_tuned'), ('mfma_fp64', 'ader_dense_f64_mfma_fp64')):
      _, kd = catalog.load(name)
      bufs_pd['Qp'] = torch.zeros_like(bufs['Qp'])
      t_d = time_kernel(kd, bufs_pd, n, args.steps, args.warmup)
      got_d = bufs_pd['Qp'].cpu().numpy().reshape(n, -1)
      diff = float((np.abs(got_d - pd_out).max(axis=1) /
                    np.maximum(np.abs(pd_out).max(axis=1), 1e-300)).max())
      dense_forms[label] = {'value': n / t_d, 'ms_per_step': t_d * 1e3,
                            'over_plain_dense_k': t_pd / t_d,
                            'max_normwise_rel_diff_vs_plain_dense_k': diff,
                            'schedule': kd.generator.get_emitted().schedule}
    del bufs_pd

  # the headline kernel's result (one untimed launch), compared with the timed launches' result
  from tensorforge_amd.device import launch as _launch
  _launch(tuned, bufs, n)
  torch.cuda.synchronize()
  gpu_out = bufs['Qp'].cpu().numpy()

  if 'f32' in legs:
    e32, k32 = catalog.load('ader_f32_tuned')
    b32 = make_inputs(e32, n, device, torch.float32)
    t32 = time_kernel(k32, b32, n, args.steps, args.warmup)
    bpe32 = k32.generator.get_emitted().bytes_per_element
    o32 = b32['Qp'].double().cpu().numpy().reshape(n, -1)
    ref = gpu_out.reshape(n, -1)
    err32 = float((np.abs(o32 - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-300)).max())
    out['f32'] = {'value': n / t32, 'ms_per_step': t32 * 1e3,
                  'hbm_GBps': (bpe32 * n + k32.generator.get_emitted().shared_bytes) / t32 / 1e9,
                  'max_normwise_rel_err_vs_f64_kernel': err32}
    del b32

  if 'mixed' in legs:
    out['mixed_orders'] = mixed_leg(device, args.steps, args.warmup)

  if 'precision' in legs:
    out['precision_order7'] = precision_leg(device, args.steps, args.warmup)

  if 'flux' in legs:
    out['flux'] = flux_leg(device, args.steps, args.warmup)
  if 'deriv' in legs:
    out['derivative'] = deriv_leg(device, args.steps, args.warmup)

  times = launch_times(tuned, bufs, n, args.steps, args.warmup)
  t_tuned = median(times)
  headline_out = bufs['Qp'].cpu().numpy()
  value = n / t_tuned
  achieved = bytes_per_launch / t_tuned / 1e9
  timing = {'statistic': 'median of per-launch HIP-event durations on the launch stream',
            'median_ms': t_tuned * 1e3, 'mean_ms': sum(times) / len(times) * 1e3,
            'min_ms': min(times) * 1e3, 'max_ms': max(times) * 1e3}

  if not np.array_equal(headline_out, gpu_out):
    raise AssertionError('the timed launches gave a different result than the first launch')
  if 'plain' in legs:
    got = gpu_out.reshape(n, -1)
    agree = float((np.abs(got - plain_out).max(axis=1) /
                   np.maximum(np.abs(plain_out).max(axis=1), 1e-300)).max())
    out['plain'] = {'value': n / t_plain, 'ms_per_step': t_plain * 1e3,
                    'tuned_over_plain': t_plain / t_tuned,
                    'max_normwise_rel_diff_tuned_vs_plain': agree}
    out['plain_dense_k'] = {'value': n / t_pd, 'ms_per_step': t_pd * 1e3,
                            'tuned_over_plain_dense_k': t_pd / t_tuned,
                            'dense_k_forms': dense_forms}

  if 'pointwise' in legs:
    out['pointwise'] = pointwise_leg(device, args.steps, args.warmup)

  if 'gemm' in legs:
    out['single_gemm'] = gemm_leg(device, args.steps, args.warmup, min(args.cpu_seconds, 8.0),
                                  not args.no_cpu)

  if 'tuner' in legs:
    out['cost_model'] = tuner_leg()

  cpu = None
  if not args.no_cpu:
    cpu = cpu_baseline(entry, bufs, gpu_out, args.cpu_seconds, full=args.cpu_full)

  line = {
    'metric': 'elements/s + % HBM roofline, fused ADER chain, tuned vs plain at stated rel err',
    'value': value,
    'unit': 'elements/s',
    'n_gpus': 1,
    'steps': args.steps,
    'warmup': args.warmup,
    'ms_per_step': t_tuned * 1e3,
    'timing': timing,
    'higher_is_better': True,
    'scaling': 'weak',
    'vs_baseline': None,
    'dtype': 'f64',
    'data': 'synthetic (seeded: Q ~ N(0,1), S star U[-2,2], K CSR U[-1,1])',
    'config': {'workload': 'fused ADER-DG volume chain Q\'=sum_d K_d Q S_d, order 6 '
                           '(56x9), 262144 elements, K_d 56x56 ~40% sparse CSR, S_d 9x9 star '
                           '(24 nnz), FP64', 'elements': n, 'emitter': 'tuned',
               'schedule': emitted.schedule},
    'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBPS, 'unit': 'GB/s',
                 'frac': achieved / HBM_PEAK_GBPS,
                 'traffic': profiled_traffic(f'{entry.name}_{n}', emitted.schedule),
                 'algorithmic_bytes_per_launch': bytes_per_launch,
                 'bytes_per_element': emitted.bytes_per_element,
                 'measured_copy_GBps': measured_bw,
                 'frac_of_measured_copy': (achieved / measured_bw) if measured_bw else None,
                 'fp64_tflops': flops / t_tuned / 1e12,
                 'fp64_frac': flops / t_tuned / 1e12 / FP64_PEAK_TFLOPS},
  }
  line.update(out)
  if cpu is not None:
    line['cpu_baseline'] = {
      'value': cpu['team']['value'], 'unit': 'elements/s', 'cores': cpu['cores'], 'kind': 'port',
      'sample': (f"first {cpu['team']['elements']} of {n} elements of the same inputs, OpenMP "
                 f"team form (reference omptarget structure: 32 lanes x 8 elements per team), "
                 f"{cpu['team']['seconds']:.1f} s"),
      'host': cpu['host'],
      'flat_form': cpu['flat'],
      'generated_omptarget': cpu['generated'],
      'gpu_max_normwise_rel_err_vs_cpu': cpu.get('gpu_max_normwise_rel_err')}
  print(json.dumps(line))


if __name__ == '__main__':
  main()
