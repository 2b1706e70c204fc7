"""Sparsity patterns: which entries of a tensor are stored, and where in the packed value array.

Same classes as the reference (``tensorforge/common/matrix/spp.py:4-109``). ``linear_index`` of a
stored entry gives its position in the caller's value array:

* ``FullSPP``: column-major over the full shape;
* ``BoundingBoxSPP``: column-major over the box;
* ``MaskSPP``: stored entries numbered in column-major order of the mask;
* ``ListSPP``: stored entries numbered in the order of the given list.
"""
from functools import reduce

import numpy as np


class SparsityPattern:
  @classmethod
  def from_values(cls, data, threshold=1):
    raise NotImplementedError

  def nonzeros(self):
    """All stored index tuples, ordered by their packed position."""
    raise NotImplementedError


def _colmajor_index(index, sizes, lower=None):
  lin, stride = 0, 1
  for d, (x, s) in enumerate(zip(index, sizes)):
    lin += (x - (lower[d] if lower is not None else 0)) * stride
    stride *= s
  return lin


class FullSPP(SparsityPattern):
  def __init__(self, shape, permute=None):
    self.shape = tuple(shape)

  def is_nz(self, index):
    return True

  def count_nz(self):
    return reduce(lambda x, y: x * y, self.shape, 1)

  def linear_index(self, tupleindex):
    return _colmajor_index(tupleindex, self.shape)

  def dimrange(self, dim):
    return (0, self.shape[dim])

  def nonzeros(self):
    return [tuple(reversed(idx)) for idx in np.ndindex(*self.shape[::-1])]


class BoundingBoxSPP(SparsityPattern):
  def __init__(self, bbox):
    self.bbox = bbox

  def is_nz(self, index):
    return self.bbox.contains(index)

  def count_nz(self):
    return reduce(lambda x, y: x * y, self.bbox.sizes(), 1)

  def linear_index(self, tupleindex):
    for x, lo, up in zip(tupleindex, self.bbox.lower(), self.bbox.upper()):
      assert lo <= x < up
    return _colmajor_index(tupleindex, self.bbox.sizes(), self.bbox.lower())

  def dimrange(self, dim):
    return (self.bbox.lower()[dim], self.bbox.upper()[dim])

  def nonzeros(self):
    lo = self.bbox.lower()
    return [tuple(i + l for i, l in zip(reversed(idx), lo))
            for idx in np.ndindex(*self.bbox.sizes()[::-1])]


class _IndexMaskSPP(SparsityPattern):
  """Shared part of ``MaskSPP``/``ListSPP``: an int mask holding packed position + 1."""

  def is_nz(self, index):
    return bool(self.indexmask[tuple(index)] > 0)

  def count_nz(self):
    return int(np.count_nonzero(self.indexmask))

  def linear_index(self, tupleindex):
    tupleindex = tuple(tupleindex)
    if len(tupleindex) == self.indexmask.ndim and \
        all(0 <= t < s for t, s in zip(tupleindex, self.indexmask.shape)):
      pos = int(self.indexmask[tupleindex])
      return pos - 1 if pos > 0 else None
    return None

  def dimrange(self, dim):
    return None

  def nonzeros(self):
    where = np.argwhere(self.indexmask > 0)
    order = np.argsort([self.indexmask[tuple(w)] for w in where], kind='stable')
    return [tuple(int(v) for v in where[o]) for o in order]


class MaskSPP(_IndexMaskSPP):
  def __init__(self, mask):
    self.mask = np.asarray(mask)
    self.indexmask = np.zeros(self.mask.shape, dtype=np.int64, order='F')
    flat_mask = self.mask.reshape(-1, order='F').astype(bool)
    positions = np.cumsum(flat_mask) * flat_mask
    self.indexmask[...] = positions.reshape(self.mask.shape, order='F')


class ListSPP(_IndexMaskSPP):
  """Entries stored in list order. ``shape`` may be omitted (the yateto frontend calls it with the
  entry list only, ``frontend/yateto.py:245``); it is then the smallest box holding the entries."""

  def __init__(self, lst, shape=None):
    self.list = [tuple(int(x) for x in entry) for entry in lst]
    if shape is None:
      if not self.list:
        raise ValueError('ListSPP needs a shape when the entry list is empty')
      shape = tuple(max(e[d] for e in self.list) + 1 for d in range(len(self.list[0])))
    self.indexmask = np.zeros(tuple(shape), dtype=np.int64, order='F')
    for pos, entry in enumerate(self.list):
      self.indexmask[entry] = pos + 1
