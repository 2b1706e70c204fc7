"""Calling generated launchers on torch device tensors.

torch only provides device memory and streams here; the arithmetic is in the generated HIP code.
``launch`` assembles the launcher's positional arguments (reference ABI) from tensors keyed by
the tensors' aliases.
"""
from typing import Dict, Optional

from .common.basic_types import Addressing


def pointer_table(tensors, device):
  """A device array holding ``tensors[e].data_ptr()``: the ``T**`` of a PTR_BASED operand."""
  import torch
  return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64, device=device)


def launch_args(generator, buffers: Dict, num_elements: int, offsets: Optional[Dict] = None,
                flags=None, stream=None, keepalive=None):
  """Positional launcher arguments.

  ``buffers[alias]``: a device tensor (NONE / STRIDED), a device ``int64`` pointer table or a
  list of device tensors (PTR_BASED), or a number (runtime SCALAR).
  """
  import torch
  offsets = offsets or {}
  args = []
  for tensor in generator._global_tensors:
    key = tensor.alias if tensor.alias is not None else tensor.name
    if tensor.addressing == Addressing.SCALAR and tensor.data is not None:
      args.append(float(tensor.data[0]) if not isinstance(tensor.data, dict) else 0.0)
      continue
    if tensor.addressing == Addressing.SCALAR:
      integral = tensor.datatype is not None and tensor.datatype.name in (
        'BOOL', 'INT', 'UINT', 'LONG', 'ULONG')
      args.append(int(buffers[key]) if integral else float(buffers[key]))
      continue
    if tensor.has_values() and len(tensor.get_values()) < 16:
      args.append(0)   # baked constant: the pointer is not read
      continue
    buf = buffers[key]
    if tensor.addressing == Addressing.PTR_BASED and isinstance(buf, (list, tuple)):
      table = pointer_table(buf, buf[0].device)
      if keepalive is not None:
        keepalive.append(table)
      buf = table
    args.append(buf.data_ptr())
    if tensor.addressing != Addressing.NONE:
      args.append(int(offsets.get(key, 0)))
  args.append(int(num_elements))
  args.append(0 if flags is None else flags.data_ptr())
  if stream is None:
    stream = torch.cuda.current_stream()
  args.append(stream.cuda_stream if hasattr(stream, 'cuda_stream') else int(stream))
  return args


def launch(kernel, buffers, num_elements, offsets=None, flags=None, stream=None):
  keep = []
  args = launch_args(kernel.generator, buffers, num_elements, offsets, flags, stream, keep)
  kernel(*args)
  return keep
