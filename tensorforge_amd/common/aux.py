"""Helpers for chain temporaries (reference ``tensorforge/common/aux.py:7-34``)."""
from typing import List

from .basic_types import Addressing, GeneralLexicon
from .matrix.tensor import Tensor


def generate_tmp_tensor(ops: List, target: List[List[int]], alias=None):
  """Shape of ``Π ops`` under ``target``: output dim ``t`` takes the extent of whichever operand
  dim maps to ``t`` (operand extents are bounding-box sizes)."""
  rank = 0
  for itarget in target:
    for t in itarget:
      rank = max(t + 1, rank)
  shape = [0] * rank
  for op, itarget in zip(ops, target):
    for j, t in enumerate(itarget):
      if t >= 0:
        shape[t] = op.bbox.sizes()[j]
  return Tensor(shape=shape, addressing=Addressing.PTR_BASED, bbox=None, is_tmp=True, alias=alias)


def generate_tmp_matrix(op1, op2, trans_a: bool = False, trans_b: bool = False):
  target_a = [-1, 0] if trans_a else [0, -1]
  target_b = [1, -1] if trans_b else [-1, 1]
  return generate_tmp_tensor([op1, op2], [target_a, target_b])


def get_extra_offset_name(symbol):
  return f'{symbol.name}{GeneralLexicon.EXTRA_OFFSET}'
