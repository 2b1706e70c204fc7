"""Grouped persistent launch over mixed orders and HIP-graph replay of time-step loops (GPU).

Members: the volume chain with dense K_d at basis sizes 4/10/20/35/56/84 (orders 2..7). Each
member's tuned and plain launchers are checked against the FP64 oracle; the grouped launcher
must reproduce the per-member launches; a captured sequence of steps with the degrees of
freedom ping-ponging between two arrays must reproduce the same steps launched one by one, on
every replay.
"""
import numpy as np
import pytest

from harness import TOL, key_of, make_inputs, normwise_error, oracle_outputs

pytestmark = pytest.mark.gpu

COUNTS = [9, 14, 21, 8, 15, 7]      # partial element groups in every member


@pytest.fixture(scope='module')
def mixed():
  from tensorforge_amd import catalog
  return catalog.load_group('ader_mixed_f64')


def _device_bufs(host, device):
  import torch
  return {k: torch.from_numpy(v.copy()).to(device) for k, v in host.items()}


def _member_inputs(group, counts, seed=21):
  return [make_inputs(e, n, seed=seed + i) for i, (e, n) in enumerate(zip(group.members, counts))]


def _grouped_args(group, grouped, bufs, counts, stream=None):
  from tensorforge_amd.device import launch_args
  lists = [launch_args(e.generator, b, n, stream=stream)
           for e, b, n in zip(group.members, bufs, counts)]
  return group.grouped.member_args(lists)


def test_members_match_oracle(mixed, device):
  import torch
  from tensorforge_amd.device import launch
  group, tuned, plain, _ = mixed
  hosts = _member_inputs(group, COUNTS)
  for e, kt, kp, host, n in zip(group.members, tuned, plain, hosts, COUNTS):
    ref = oracle_outputs(e, host, n)['Qp']
    vol = e.tensors['Qp'].get_real_volume()
    for k in (kt, kp):
      bufs = _device_bufs(host, device)
      launch(k, bufs, n)
      torch.cuda.synchronize()
      err = normwise_error(bufs['Qp'].cpu().numpy(), ref, vol)
      print(f'b={e.meta["b"]} n={n} {k.name}: {err:.2e}')
      assert err <= TOL[e.fp_type], (e.meta['b'], k.name, err)


@pytest.mark.parametrize('counts', [COUNTS, [0, 14, 0, 8, 15, 0], [300, 1, 64, 63, 65, 129]])
def test_grouped_equals_member_launches(mixed, device, counts):
  import torch
  from tensorforge_amd.device import launch
  group, tuned, _, gk = mixed
  hosts = _member_inputs(group, counts, seed=5)
  solo = []
  for k, host, n in zip(tuned, hosts, counts):
    bufs = _device_bufs(host, device)
    launch(k, bufs, n)
    solo.append(bufs)
  together = [_device_bufs(h, device) for h in hosts]
  gk(*_grouped_args(group, gk, together, counts))
  torch.cuda.synchronize()
  for e, a, b, n, host in zip(group.members, solo, together, counts, hosts):
    if n == 0:
      continue
    vol = e.tensors['Qp'].get_real_volume()
    err = normwise_error(b['Qp'].cpu().numpy(), a['Qp'].cpu().numpy(), vol)
    ref = oracle_outputs(e, host, n)['Qp']
    err_ref = normwise_error(b['Qp'].cpu().numpy(), ref, vol)
    print(f'b={e.meta["b"]} n={n}: grouped vs member {err:.2e}, vs oracle {err_ref:.2e}')
    assert err <= 1e-14 and err_ref <= TOL[e.fp_type]


def test_graph_replay_of_ping_pong_steps(mixed, device):
  """Four steps Q -> Q' -> Q -> ... captured once; three replays, with eager launches and
  synchronisations in between, each equal to the eager sequence."""
  import torch
  from tensorforge_amd.graph import Graph
  group, _, _, gk = mixed
  counts = [5, 9, 12, 7, 10, 6]
  steps = 4
  hosts = _member_inputs(group, counts, seed=9)
  stream = torch.cuda.Stream()

  def run_steps(bufs, launch_stream):
    for s in range(steps):
      cur = [dict(b, Q=b['Q'], Qp=b['Qp']) if s % 2 == 0 else dict(b, Q=b['Qp'], Qp=b['Q'])
             for b in bufs]
      gk(*_grouped_args(group, gk, cur, counts, stream=launch_stream))

  eager = [_device_bufs(h, device) for h in hosts]
  with torch.cuda.stream(stream):
    run_steps(eager, stream)
  torch.cuda.synchronize()
  final_key = 'Q' if steps % 2 == 0 else 'Qp'
  want = [b[final_key].cpu().numpy() for b in eager]

  work = [_device_bufs(h, device) for h in hosts]
  initial = [{k: v.clone() for k, v in b.items()} for b in work]
  graph = Graph(stream)
  # one launch outside capture: the launcher's one-time device queries
  scratch = [_device_bufs(h, device) for h in hosts]
  gk(*_grouped_args(group, gk, scratch, counts, stream=stream))
  torch.cuda.synchronize()
  graph.capture(lambda: run_steps(work, stream))
  for replay in range(3):
    for b, init in zip(work, initial):
      for k in b:
        b[k].copy_(init[k])
    torch.cuda.synchronize()
    graph.launch()
    torch.cuda.synchronize()
    # eager work between replays
    gk(*_grouped_args(group, gk, scratch, counts))
    torch.cuda.synchronize()
    for e, b, w in zip(group.members, work, want):
      got = b[final_key].cpu().numpy()
      vol = e.tensors['Q'].get_real_volume()
      err = normwise_error(got, w, vol)
      assert err <= 1e-14, (replay, e.meta['b'], err)
  graph.close()


def test_many_rounds_grouped_and_members_match_plain(mixed, device):
  """Counts that give every member several element groups per workgroup: the grouped launch and
  each member's tuned launcher (autotuned schedules) against the member's plain launcher, which
  has no cross-group state, on every element."""
  import torch
  from tensorforge_amd.device import launch
  group, tuned, plain, gk = mixed
  counts = [30011, 60013, 40007, 35003, 30001, 30007]
  print(f'grouped and tuned members vs plain, counts {counts}', flush=True)
  hosts = _member_inputs(group, counts, seed=31)
  ref, solo = [], []
  for kp, kt, host, n in zip(plain, tuned, hosts, counts):
    bufs = _device_bufs(host, device)
    launch(kp, bufs, n)
    ref.append(bufs['Qp'].cpu().numpy())
    bufs = _device_bufs(host, device)
    launch(kt, bufs, n)
    solo.append(bufs['Qp'].cpu().numpy())
    torch.cuda.synchronize()
  together = [_device_bufs(h, device) for h in hosts]
  gk(*_grouped_args(group, gk, together, counts))
  torch.cuda.synchronize()
  for e, r, s, t, n in zip(group.members, ref, solo, together, counts):
    vol = e.tensors['Qp'].get_real_volume()
    err_solo = normwise_error(s, r, vol)
    err_group = normwise_error(t['Qp'].cpu().numpy(), r, vol)
    print(f'b={e.meta["b"]} n={n}: tuned member vs plain {err_solo:.2e}, grouped vs plain '
          f'{err_group:.2e}', flush=True)
    assert err_solo <= 1e-12 and err_group <= 1e-12
