"""Shared test plumbing: seeded inputs for a catalog entry, the oracle run, the GPU run, and
the error measure.

Error measure (stated with every result): normwise per element,
``max_e ||x_e - ref_e||_inf / ||ref_e||_inf``.
"""
import numpy as np

from oracle.chain_oracle import OracleInputs, run as oracle_run
from tensorforge_amd.common.basic_types import Addressing, DataFlowDirection, FloatingPointType

TOL = {FloatingPointType.DOUBLE: 1e-12, FloatingPointType.FLOAT: 1e-5}


def np_dtype(fp):
  return np.float64 if fp == FloatingPointType.DOUBLE else np.float32


def make_inputs(entry, num_elements, seed=0, offsets=None, garbage_unused=True):
  """Host inputs per tensor alias, in the launcher's layout.

  Unused slots of packed sparse tensors hold random values so a kernel that reads them fails."""
  rng = np.random.default_rng(seed)
  gen = entry.generator
  dt = np_dtype(entry.fp_type)
  offsets = dict(offsets or {})
  host = {}
  for t in gen._global_tensors:
    key = key_of(t)
    vol = t.get_real_volume()
    if t.addressing == Addressing.SCALAR:
      if t.data is None:
        host[key] = float(rng.uniform(0.5, 2.0))
      continue
    if t.has_values() and len(t.get_values()) < 16:
      continue
    off = int(offsets.get(key, 0))
    if t.addressing == Addressing.NONE:
      host[key] = rng.uniform(-1, 1, size=vol).astype(dt)
    elif t.addressing == Addressing.STRIDED:
      host[key] = rng.uniform(-1, 1, size=num_elements * vol + off).astype(dt)
    else:
      host[key] = [rng.uniform(-1, 1, size=vol + off).astype(dt) for _ in range(num_elements)]
  return host


def oracle_outputs(entry, host, num_elements, offsets=None, flags=None):
  gen = entry.generator
  arrays = {}
  for t in gen._global_tensors:
    if key_of(t) in host:
      v = host[key_of(t)]
      arrays[t] = [x.astype(np.float64).copy() for x in v] if isinstance(v, list) else \
        (np.asarray(v, dtype=np.float64).copy() if not isinstance(v, float) else v)
  offs = {t: (offsets or {}).get(key_of(t), 0) for t in gen._global_tensors}
  inp = OracleInputs(arrays, num_elements, flags=flags, offsets=offs)
  out = oracle_run(entry.descrs, inp)
  return {key_of(t): out[t] for t in gen._global_tensors if t in out}


def key_of(t):
  """Buffer key of a tensor: its alias, else its generated name."""
  return t.alias if t.alias is not None else t.name


def outputs_of(entry):
  return [key_of(t) for t in entry.generator._global_tensors
          if t.direction in (DataFlowDirection.SINK, DataFlowDirection.SOURCESINK)]


def gpu_run(kernel, entry, host, num_elements, device, offsets=None, flags=None):
  import torch
  from tensorforge_amd.device import launch
  bufs = {}
  for key, v in host.items():
    if isinstance(v, list):
      bufs[key] = [torch.from_numpy(x.copy()).to(device) for x in v]
    elif isinstance(v, float):
      bufs[key] = v
    else:
      bufs[key] = torch.from_numpy(v.copy()).to(device)
  flags_t = None if flags is None else torch.from_numpy(np.asarray(flags, dtype=np.uint32)).to(device)
  keep = launch(kernel, bufs, num_elements, offsets=offsets, flags=flags_t)
  torch.cuda.synchronize()
  del keep
  out = {}
  for key, v in bufs.items():
    if isinstance(v, list):
      out[key] = [x.cpu().numpy() for x in v]
    elif not isinstance(v, float):
      out[key] = v.cpu().numpy()
  return out


def normwise_error(got, ref, vol, offset=0):
  """max over elements of ||got_e - ref_e||_inf / ||ref_e||_inf."""
  if isinstance(ref, list):
    g = np.stack([np.asarray(x, dtype=np.float64) for x in got])
    r = np.stack([np.asarray(x, dtype=np.float64) for x in ref])
  else:
    n = (len(ref) - offset) // vol * vol
    g = np.asarray(got[offset:offset + n], dtype=np.float64).reshape(-1, vol)
    r = np.asarray(ref[offset:offset + n], dtype=np.float64).reshape(-1, vol)
  num = np.abs(g - r).max(axis=1)
  den = np.maximum(np.abs(r).max(axis=1), np.finfo(np.float64).tiny)
  return float((num / den).max()) if len(num) else 0.0
