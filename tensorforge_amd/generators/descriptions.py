"""Contraction descriptions: the public API a caller builds kernels from.

Same classes and constructor signatures as the reference
(``tensorforge/generators/descriptions.py:14-149``).

Semantics of ``MultilinearDescr(dest, ops, target, permute, add)`` as implemented here:

* operand ``i``'s logical dimension ``j`` is its storage dimension ``permute[i][j]``
  (identity when ``permute[i]`` is empty or ``None``);
* ``target[i][j] >= 0`` maps logical dim ``j`` to output dim ``target[i][j]``;
  ``target[i][j] < 0`` to the contracted index ``-target[i][j] - 1``;
* ``dest[n] = alpha * sum_k prod_i op_i[...]  (+ beta * dest[n] when add)``.

Deviations from the reference's emitted arithmetic, each deliberate:

* The reference's loader stores a permuted operand with the same indices it read it with
  (``backend/instructions/memory/load.py:143-154``), so ``permute`` never reaches the arithmetic
  and ``GemmDescr(trans_a=True)`` computes ``A·B``. Here ``permute`` is applied, so ``trans_a``
  means ``Aᵀ``.
* ``GemmDescr`` in the reference forwards ``strict_match`` into the ``add`` slot and drops
  ``beta`` (``descriptions.py:146``), and a numeric ``alpha`` is read back as absent
  (``common/matrix/tensor.py:61-66``). Here ``C = alpha·op(A)·op(B) + beta·C`` as the reference's
  own alpha/beta test suites expect (``tests/gemms/testsuites/alpha-beta-tests.yaml``).
  A symbolic ``beta`` has no launcher argument in the reference ABI, so it is taken as 1.
"""
from typing import List

from ..common.basic_types import Addressing, DataFlowDirection, FloatingPointType
from ..common.context import Context
from ..common.exceptions import GenerationError
from ..common.matrix.tensor import SubTensor, Tensor


class OperationDescription:
  pass


class MultilinearDescr(OperationDescription):
  def __init__(self, dest: SubTensor, ops: List[SubTensor], target, permute, add: bool = False,
               strict_match: bool = False, prefer_align: bool = False):
    self.dest = dest
    self.ops = ops
    self.target = target
    self.permute = permute
    self._strict_match = False
    self.add = add
    self.beta = 1.0 if add else 0.0
    self.prefer_align = prefer_align

    if len(target) != len(ops):
      raise GenerationError(f'{len(ops)} operands but {len(target)} target maps')
    self.dest.tensor.set_data_flow_direction(DataFlowDirection.SINK)
    for op in self.ops:
      op.tensor.set_data_flow_direction(DataFlowDirection.SOURCE)

  def _lead_dim(self):
    return self.dest.bbox.sizes()[0]

  def get_num_threads(self, context: Context):
    """Lanes per element, the reference heuristic (``descriptions.py:35-49``): the leading
    extent aligned to a wavefront's words, then snapped down to a power of two <= 32 when the
    extent is that small."""
    lead = self._lead_dim()
    num_threads = context.align(num=lead)
    for cap in (32, 16, 8, 4, 2, 1):
      if lead <= cap:
        num_threads = cap
    return num_threads, lead

  def get_accumulator_size(self):
    acc = 1
    for s in self.dest.bbox.sizes()[1:]:
      acc *= s
    return acc

  def is_strict_match(self):
    return self._strict_match

  def matrix_list(self):
    return [self.dest] + list(self.ops)

  def logical_permute(self, i):
    """Permutation of operand ``i`` (identity when not given)."""
    rank = len(self.target[i])
    perm = self.permute[i] if self.permute is not None and i < len(self.permute) else None
    if perm is None or len(perm) == 0:
      return list(range(rank))
    if sorted(perm) != list(range(rank)):
      raise GenerationError(f'permute {perm} of operand {i} is not a permutation of {rank} dims')
    return list(perm)

  def __str__(self):
    desttarget = [i for i in range(self.dest.bbox.rank())]
    return (f'{self.dest}{desttarget} = '
            f'{"×".join(f"{op}{optarget}" for op, optarget in zip(self.ops, self.target))}')


class ElementwiseDescr(OperationDescription):
  """Pointwise statements over every index point of every element (reference
  ``descriptions.py:67-95``): ``oplist`` is a list of ``optree`` statements (``Assignment``,
  ``IfNode``, ``WhileNode``). Tensors assigned to become sinks, tensors read become sources
  (both: source-sink). Temporaries without a name are numbered ``tmp0, tmp1, ...`` in
  first-appearance order, so ``str()`` (and the kernel name) is deterministic."""

  def __init__(self, oplist, strict_match: bool = False, prefer_align: bool = False):
    from . import optree
    self.oplist = list(oplist)
    self._strict_match = False
    for op in self.oplist:
      if not isinstance(op, optree.Statement):
        raise GenerationError(f'ElementwiseDescr takes optree statements, got {type(op).__name__}')
      for sub in op.tensors(outtensors=True, intensors=False):
        sub.tensor.set_data_flow_direction(DataFlowDirection.SINK)
      for sub in op.tensors(outtensors=False, intensors=True):
        sub.tensor.set_data_flow_direction(DataFlowDirection.SOURCE)
    count = 0
    for s in optree.walk_statements(self.oplist):
      for root in optree.statement_nodes(s):
        for node in optree.walk_nodes(root):
          if isinstance(node, optree.TempVar) and node.name is None:
            node.name = f'tmp{count}'
            count += 1

  def get_num_threads(self, context: Context):
    return 64, 64

  def get_accumulator_size(self):
    return 0

  def is_strict_match(self):
    return self._strict_match

  def matrix_list(self):
    return [tensor for op in self.oplist for tensor in op.tensors()]

  def __str__(self):
    return '; '.join(str(op) for op in self.oplist)


class GemmDescr(MultilinearDescr):
  """``C = alpha·op(A)·op(B) + beta·C`` on rank-2 operands."""

  def __init__(self, trans_a, trans_b, a, b, c, alpha=1.0, beta=0.0, strict_match: bool = False,
               prefer_align: bool = False):
    target_a = [0, -1]
    target_b = [-1, 1]
    permute_a = [1, 0] if trans_a else [0, 1]
    permute_b = [1, 0] if trans_b else [0, 1]
    self.trans_a = trans_a
    self.trans_b = trans_b
    self.alpha = alpha
    if isinstance(beta, str):
      beta_value = 1.0
    else:
      beta_value = float(beta)
    if alpha == 1.0:
      super().__init__(c, [a, b], [target_a, target_b], [permute_a, permute_b],
                       add=beta_value != 0.0, prefer_align=prefer_align)
    else:
      is_number = isinstance(alpha, (float, int))
      alpha_tensor = SubTensor(Tensor([], Addressing.SCALAR,
                                      data=[float(alpha)] if is_number else None))
      super().__init__(c, [a, b, alpha_tensor], [target_a, target_b, []],
                       [permute_a, permute_b, []], add=beta_value != 0.0,
                       prefer_align=prefer_align)
    self.beta = beta_value
