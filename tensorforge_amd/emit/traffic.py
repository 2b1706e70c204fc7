"""Compulsory memory traffic of a chain program: the algorithmic bytes a roofline is priced on.

Per element: every entry of a per-element operand that the chain reads before writing it, counted
once however often the chain uses it (packed nonzeros only for sparse operands), plus every entry
it writes. Operands shared by all elements (``Addressing.NONE``) are counted once per launch.
"""
from dataclasses import dataclass

from . import ir


@dataclass
class Traffic:
  per_element: int      # bytes per element
  shared: int           # bytes per launch
  read_per_element: int
  write_per_element: int


def _factor_offsets(p, f):
  window = [p.n_ranges[i] if kind == 'n' else p.k_ranges[i] for kind, i in f.labels]
  if any(hi <= lo for lo, hi in window):
    return set()
  if f.is_sparse():
    return {pos for idx, pos in f.packed.items()
            if all(lo <= x < hi for x, (lo, hi) in zip(idx, window))}
  return {f.offset(idx) for idx in ir.iterate_box(window)}


def element_traffic(program: ir.ChainProgram) -> Traffic:
  read, written = {}, {}
  for p in program.products:
    if p.out_empty():
      continue
    for f in p.factors:
      if f.src != ir.SRC_GLOBAL:
        continue
      offs = _factor_offsets(p, f)
      key = f.arg.name
      w = written.setdefault(key, set())
      read.setdefault(key, set()).update(o for o in offs if o not in w)
    if p.dest_src == ir.SRC_GLOBAL:
      key = p.dest_arg.name
      offs = {p.dest_offset(n) for n in ir.iterate_box(p.n_ranges)}
      w = written.setdefault(key, set())
      if p.add:
        read.setdefault(key, set()).update(o for o in offs if o not in w)
      w.update(offs)
  per_r = per_w = shared = 0
  for a in program.args:
    size = a.dtype.size()
    r = len(read.get(a.name, ())) * size
    wr = len(written.get(a.name, ())) * size
    if a.per_element:
      per_r += r
      per_w += wr
    else:
      shared += r + wr
  return Traffic(per_element=per_r + per_w, shared=shared, read_per_element=per_r,
                 write_per_element=per_w)
