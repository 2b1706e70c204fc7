"""Timing of generated kernels on synthetic inputs of the right shapes (device memory via torch).

Values are random: timing does not depend on them except through denormals, which uniform
[-1, 1] data does not produce.
"""
import ctypes

from ..common.basic_types import Addressing


def synthetic_buffers(generator, num_elements, device, seed=0):
  """Device buffers for every launcher operand: ``{key: tensor | float}``."""
  import torch
  g = torch.Generator(device=device)
  g.manual_seed(seed)
  dt = {'double': torch.float64, 'float': torch.float32}
  bufs = {}
  keep = []
  for t in generator._global_tensors:
    key = t.alias if t.alias is not None else t.name
    dtype = dt[str(generator._arg_type(t))]
    vol = t.get_real_volume()
    if t.addressing == Addressing.SCALAR:
      bufs[key] = 1.0
      continue
    if t.has_values() and len(t.get_values()) < 16:
      continue
    if t.addressing == Addressing.NONE:
      bufs[key] = torch.rand(vol, generator=g, device=device, dtype=dtype) * 2 - 1
    elif t.addressing == Addressing.STRIDED:
      bufs[key] = torch.rand(num_elements * vol, generator=g, device=device, dtype=dtype) * 2 - 1
    else:
      big = torch.rand(num_elements * vol, generator=g, device=device, dtype=dtype) * 2 - 1
      keep.append(big)
      base = big.data_ptr()
      size = big.element_size()
      bufs[key] = torch.tensor([base + e * vol * size for e in range(num_elements)],
                               dtype=torch.int64, device=device)
  bufs['__keep__'] = keep
  return bufs


def time_launcher(kernel, bufs, num_elements, steps=10, warmup=2):
  """Average duration (s) of ``steps`` launches, HIP events on the launch stream."""
  import torch
  from ..device import launch_args
  stream = torch.cuda.current_stream()
  clean = {k: v for k, v in bufs.items() if k != '__keep__'}
  args = launch_args(kernel.generator, clean, num_elements, stream=stream)
  for _ in range(warmup):
    kernel(*args)
  torch.cuda.synchronize()
  start = torch.cuda.Event(enable_timing=True)
  stop = torch.cuda.Event(enable_timing=True)
  start.record(stream)
  for _ in range(steps):
    kernel(*args)
  stop.record(stream)
  torch.cuda.synchronize()
  return start.elapsed_time(stop) / 1e3 / steps
