"""Replaying launcher sequences as HIP graphs (``tensorforge_graph_*`` in the runtime library).

A time-step loop of generated launches (for example ten steps of the volume chain with the
degrees of freedom ping-ponging between two arrays) is captured once and replayed with one
``hipGraphLaunch`` instead of one launch call per kernel.
"""
import ctypes

from .runtime.loader import runtime


def _check(code, what):
  if code != 0:
    rt = runtime()
    msg = rt.tensorforge_error_string(code).decode()
    rt.tensorforge_clear_error()
    raise RuntimeError(f'{what}: HIP error {code} ({msg})')


class Graph:
  """``Graph(stream).capture(fn)`` records the launches ``fn()`` issues on ``stream`` (a torch
  stream or a raw ``hipStream_t`` value, not the null stream); ``launch()`` replays them."""

  def __init__(self, stream):
    self.stream = int(stream.cuda_stream if hasattr(stream, 'cuda_stream') else stream)
    if self.stream == 0:
      raise ValueError('graphs are captured on a created stream, not the null stream')
    self.exec = ctypes.c_void_p()

  def capture(self, fn):
    rt = runtime()
    _check(rt.tensorforge_graph_begin(ctypes.c_void_p(self.stream)), 'tensorforge_graph_begin')
    try:
      fn()
    finally:
      code = rt.tensorforge_graph_end(ctypes.c_void_p(self.stream), ctypes.byref(self.exec))
    _check(code, 'tensorforge_graph_end')
    return self

  def launch(self):
    _check(runtime().tensorforge_graph_launch(self.exec, ctypes.c_void_p(self.stream)),
           'tensorforge_graph_launch')

  def close(self):
    if self.exec:
      runtime().tensorforge_graph_destroy(self.exec)
      self.exec = ctypes.c_void_p()

  def __del__(self):
    try:
      self.close()
    except Exception:   # interpreter shutdown
      pass
