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


_INT_DTYPES = {'INT': np.int32, 'UINT': np.uint32, 'LONG': np.int64, 'ULONG': np.uint64,
               'BOOL': np.bool_}


def tensor_dtype(t, fp):
  """numpy dtype of a tensor's entries: its datatype, else the kernel's floating type."""
  if t.datatype is None:
    return np_dtype(fp)
  if t.datatype.name in _INT_DTYPES:
    return _INT_DTYPES[t.datatype.name]
  return np_dtype(t.datatype)


def _draw(rng, size, dt):
  if dt == np.bool_:
    return rng.random(size) < 0.5
  if np.issubdtype(dt, np.integer):
    lo = 0 if np.issubdtype(dt, np.unsignedinteger) else -20
    return rng.integers(lo, 20, size=size).astype(dt)
  return rng.uniform(-1, 1, size=size).astype(dt)


def make_inputs(entry, num_elements, seed=0, offsets=None, garbage_unused=True):
  """Host inputs per tensor alias, in the launcher's layout.

  Unused slots of packed sparse tensors hold random values so a kernel that reads them fails.
  Integer tensors hold integers in [-20, 20), bool tensors random truth values."""
  rng = np.random.default_rng(seed)
  gen = entry.generator
  offsets = dict(offsets or {})
  host = {}
  for t in gen._global_tensors:
    key = key_of(t)
    vol = t.get_real_volume()
    dt = tensor_dtype(t, entry.fp_type)
    if t.addressing == Addressing.SCALAR:
      if t.data is None:
        host[key] = int(rng.integers(1, 5)) if np.issubdtype(dt, np.integer) else \
          float(rng.uniform(0.5, 2.0))
      continue
    if t.has_values() and len(t.get_values()) < 16:
      continue
    off = int(offsets.get(key, 0))
    if t.addressing == Addressing.NONE:
      host[key] = _draw(rng, vol, dt)
    elif t.addressing == Addressing.STRIDED:
      host[key] = _draw(rng, num_elements * vol + off, dt)
    else:
      host[key] = [_draw(rng, vol + off, dt) for _ in range(num_elements)]
  return host


def is_pointwise(entry):
  return any(type(d).__name__ == 'ElementwiseDescr' for d in entry.descrs)


def oracle_outputs(entry, host, num_elements, offsets=None, flags=None):
  gen = entry.generator
  arrays = {}
  for t in gen._global_tensors:
    if key_of(t) in host:
      v = host[key_of(t)]
      arrays[t] = [x.astype(np.float64).copy() for x in v] if isinstance(v, list) else \
        (np.asarray(v, dtype=np.float64).copy() if not isinstance(v, (float, int)) else v)
  offs = {t: (offsets or {}).get(key_of(t), 0) for t in gen._global_tensors}
  inp = OracleInputs(arrays, num_elements, flags=flags, offsets=offs)
  chain = [d for d in entry.descrs if type(d).__name__ != 'ElementwiseDescr']
  statements = [d for d in entry.descrs if type(d).__name__ == 'ElementwiseDescr']
  out = oracle_run(chain, inp) if chain else inp.arrays
  if statements:
    # a chain's fused statements run after the chain, on what it wrote
    from oracle.pointwise_oracle import run as pointwise_run
    out = pointwise_run(statements, inp)
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
    elif isinstance(v, (float, int)):
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
    elif not isinstance(v, (float, int)):
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
