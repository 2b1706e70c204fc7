"""FP64 NumPy oracle for description lists, at the level of the launcher ABI.

TEST INFRASTRUCTURE ONLY. This module is the checker the GPU kernels are compared against; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may use it. Nothing in
``tensorforge_amd`` imports it, and it shares no code with the emitters: it interprets the
description objects itself.

It restates the reference's semantics:

* index ranges: intersection of the operand bounding boxes mapped to each output / contracted
  index (``tensorforge/backend/instructions/compute/multilinear.py:58-96``);
* element addressing: STRIDED ``e * prod(shape) + extraOffset``, PTR_BASED
  ``ptr[e] + extraOffset``, NONE shared (``tensorforge/backend/instructions/ptr_manip.py:19-65``);
* entry addressing: ``sum((idx - bbox.lower) * colmajor_stride(shape))`` for dense operands
  (``tensorforge/backend/symbol.py:304-312``), the pattern's packed position for sparse ones
  (``tensorforge/backend/symbol.py:340-401``, ``tensorforge/common/matrix/spp.py:84-109``);
* accumulation: the full contraction first, then ``+ old destination`` when ``add``
  (``tensorforge/backend/instructions/compute/multilinear.py:226-247``);
* elements with ``flags[e] == 0`` and elements ``>= numElements`` are left untouched
  (``tensorforge/generators/generator.py:494-505``).

and the intended meaning where the reference's emitted code loses it (documented with the
description classes): ``permute`` reorders an operand's dimensions before ``target`` maps them,
and ``GemmDescr`` computes ``alpha·op(A)·op(B) + beta·C``.

Pinned by ``tests/test_oracle.py`` against closed forms (batched ``matmul``), hand-computed tiny
cases, brute-force loops and invariants (linearity, flags, extra offsets, pointer-based vs
strided, sparse vs densified).
"""
import string

import numpy as np


class OracleInputs:
  """Arrays as a launcher would see them, keyed by ``Tensor`` object.

  ``arrays[t]`` is a flat array (STRIDED, NONE), a list of flat arrays (PTR_BASED) or a float
  (SCALAR). ``offsets[t]`` is the extra offset (default 0).
  """

  def __init__(self, arrays, num_elements, flags=None, offsets=None):
    self.arrays = dict(arrays)
    self.num_elements = int(num_elements)
    self.flags = None if flags is None else np.asarray(flags)
    self.offsets = dict(offsets or {})


def _colmajor_strides(shape):
  out, s = [], 1
  for x in shape:
    out.append(s)
    s *= x
  return out


def _addressing_name(t):
  return str(t.addressing)


def _is_packed(t):
  from tensorforge_amd.common.matrix.spp import BoundingBoxSPP
  return t.spp.count_nz() != int(np.prod(t.shape)) and not isinstance(t.spp, BoundingBoxSPP)


def _window_local_addresses(t, window):
  """Per-entry local address (or -1 for a structural zero) over ``window`` (per storage dim
  ``(lo, hi)``), as an int array of the window's shape (C order over the dims listed)."""
  sizes = [hi - lo for lo, hi in window]
  grids = np.meshgrid(*[np.arange(lo, hi) for lo, hi in window], indexing='ij') if window else []
  if _is_packed(t):
    flat_idx = np.stack([g.reshape(-1) for g in grids], axis=1) if window else np.zeros((1, 0), int)
    loc = np.array([-1 if t.spp.linear_index(tuple(int(x) for x in row)) is None
                    else t.spp.linear_index(tuple(int(x) for x in row)) for row in flat_idx],
                   dtype=np.int64)
    return loc.reshape(sizes)
  strides = _colmajor_strides(t.shape)
  loc = np.zeros(sizes, dtype=np.int64)
  for d, g in enumerate(grids):
    loc = loc + (g - t.bbox.lower()[d]) * strides[d]
  return loc


def _read_global(t, window, inp: OracleInputs, elems):
  """Values of tensor ``t`` over ``window`` for the element indices ``elems``:
  shape ``[len(elems)] + window sizes`` (per element) or window sizes (shared)."""
  loc = _window_local_addresses(t, window)
  zero = loc < 0
  safe = np.where(zero, 0, loc)
  addr = _addressing_name(t)
  data = inp.arrays[t]
  off = int(inp.offsets.get(t, 0))
  if addr == 'none':
    vals = np.asarray(data, dtype=np.float64)[safe]
    return np.where(zero, 0.0, vals)
  if addr == 'strided':
    vol = int(np.prod(t.shape))
    base = elems.astype(np.int64) * vol + off
    vals = np.asarray(data, dtype=np.float64)[base.reshape((-1,) + (1,) * loc.ndim) + safe[None]]
    return np.where(zero[None], 0.0, vals)
  if addr == 'pointer_based':
    vals = np.stack([np.asarray(data[e], dtype=np.float64)[off + safe] for e in elems]) \
      if len(elems) else np.zeros((0,) + loc.shape)
    return np.where(zero[None], 0.0, vals)
  raise ValueError(f'cannot read {addr} tensor {t.name} as an array')


def _write_global(t, window, values, inp: OracleInputs, elems):
  loc = _window_local_addresses(t, window)
  addr = _addressing_name(t)
  off = int(inp.offsets.get(t, 0))
  if addr == 'strided':
    vol = int(np.prod(t.shape))
    base = elems.astype(np.int64) * vol + off
    inp.arrays[t][base.reshape((-1,) + (1,) * loc.ndim) + loc[None]] = values
  elif addr == 'pointer_based':
    for i, e in enumerate(elems):
      inp.arrays[t][e][off + loc] = values[i]
  else:
    raise ValueError(f'cannot write {addr} tensor {t.name}')


def _perm(descr, i):
  rank = len(descr.target[i])
  perm = descr.permute[i] if descr.permute is not None and i < len(descr.permute) else None
  return list(range(rank)) if not perm else list(perm)


def _scalar_value(t, inp):
  if t.data is not None:
    return float(t.data[0]) if not isinstance(t.data, dict) else float(next(iter(t.data.values())))
  return float(inp.arrays[t])


def run(descrs, inp: OracleInputs):
  """Execute ``descrs`` on ``inp`` (modified in place) and return ``inp.arrays``."""
  n = inp.num_elements
  elems = np.arange(n)
  if inp.flags is not None:
    elems = elems[np.asarray(inp.flags[:n]) != 0]
  tmp_vals, tmp_origin = {}, {}
  for descr in descrs:
    nr = max([t + 1 for tg in descr.target for t in tg] + [0])
    kr = max([-t for tg in descr.target for t in tg if t < 0] + [0])
    n_rng = [[-10**9, 10**9] for _ in range(nr)]
    k_rng = [[-10**9, 10**9] for _ in range(kr)]
    for i, op in enumerate(descr.ops):
      perm = _perm(descr, i)
      for j, t in enumerate(descr.target[i]):
        p = perm[j]
        rng = n_rng[t] if t >= 0 else k_rng[-t - 1]
        rng[0] = max(rng[0], op.bbox.lower()[p])
        rng[1] = min(rng[1], op.bbox.upper()[p])
    dest = descr.dest.tensor
    if dest.is_tmp and id(dest) in tmp_origin:
      o = tmp_origin[id(dest)]
      shp = tmp_vals[id(dest)].shape[1:]
      n_rng = [[max(a, oo), min(b, oo + s)] for (a, b), oo, s in zip(n_rng, o, shp)]
    if not dest.is_tmp:
      view, tb = descr.dest.bbox, dest.bbox
      n_rng = [[max(a, vl, tl), min(b, vu, tu)] for (a, b), vl, vu, tl, tu in
               zip(n_rng, view.lower(), view.upper(), tb.lower(), tb.upper())]
    n_rng = [(a, max(a, b)) for a, b in n_rng]
    k_rng = [(a, max(a, b)) for a, b in k_rng]
    letters = iter(string.ascii_letters)
    n_let = [next(letters) for _ in range(nr)]
    k_let = [next(letters) for _ in range(kr)]
    operands, subs, scale = [], [], 1.0
    for i, op in enumerate(descr.ops):
      t = op.tensor
      perm = _perm(descr, i)
      labels = [None] * len(t.shape)
      for j, tg in enumerate(descr.target[i]):
        labels[perm[j]] = n_let[tg] if tg >= 0 else k_let[-tg - 1]
      window = [n_rng[tg] if tg >= 0 else k_rng[-tg - 1] for tg in
                [descr.target[i][perm.index(p)] for p in range(len(t.shape))]]
      if t.is_tmp:
        o = tmp_origin[id(t)]
        arr = tmp_vals[id(t)]
        sl = tuple(slice(lo - oo, hi - oo) for (lo, hi), oo in zip(window, [0] * len(window)))
        operands.append(arr[(slice(None),) + sl])
        subs.append('Z' + ''.join(labels))
      elif t.addressing.name == 'SCALAR' or (t.data is not None and len(t.data) < 16
                                            and all(s == 1 for s in t.shape)):
        scale *= _scalar_value(t, inp)
      elif t.data is not None and len(t.data) < 16:
        vals = np.zeros([hi - lo for lo, hi in window])
        for idx in np.ndindex(*vals.shape):
          v = t.value(tuple(i + lo for i, (lo, _) in zip(idx, window)))
          vals[idx] = 0.0 if v is None else v
        operands.append(vals)
        subs.append(''.join(labels))
      elif str(t.addressing) == 'none':
        operands.append(_read_global(t, window, inp, elems))
        subs.append(''.join(labels))
      else:
        operands.append(_read_global(t, window, inp, elems))
        subs.append('Z' + ''.join(labels))
    out_sub = 'Z' + ''.join(n_let)
    if not any('Z' in s for s in subs):
      operands.append(np.ones(len(elems)))
      subs.append('Z')
    result = np.einsum(','.join(subs) + '->' + out_sub, *operands, optimize=True) * scale
    beta = getattr(descr, 'beta', 1.0 if descr.add else 0.0)
    if dest.is_tmp:
      if id(dest) not in tmp_origin:
        tmp_origin[id(dest)] = tuple(lo for lo, _ in n_rng)
        tmp_vals[id(dest)] = np.zeros((len(elems),) + tuple(hi - lo for lo, hi in n_rng))
      o = tmp_origin[id(dest)]
      sl = tuple(slice(lo - oo, hi - oo) for (lo, hi), oo in zip(n_rng, o))
      cur = tmp_vals[id(dest)][(slice(None),) + sl]
      tmp_vals[id(dest)][(slice(None),) + sl] = result + beta * cur if descr.add else result
    else:
      if descr.add:
        result = result + beta * _read_global(dest, n_rng, inp, elems)
      _write_global(dest, n_rng, result, inp, elems)
  return inp.arrays
