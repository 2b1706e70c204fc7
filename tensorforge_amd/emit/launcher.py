"""Launcher (C-ABI) emission shared by both emitters, and translation-unit assembly.

The launcher keeps the reference's name and parameter list (``generators/generator.py:384-436``)
and is declared ``extern "C"``. Differences from the reference launcher
(``common/vm/lexic/hip_lexic.py:21-50``): ``numElements == 0`` returns without launching, and a
launch error is recorded for ``tensorforge_last_error()`` instead of terminating the process.
"""
from dataclasses import dataclass, field
from typing import List, Tuple

from ..common.basic_types import Addressing, GeneralLexicon

HEADER = r'''#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
extern "C" void tensorforge_record_error(int code, const char* where);
'''


@dataclass
class Emitted:
  """What an emitter produced for one chain."""
  kernel_name: str
  kernel_src: str
  block: Tuple[int, int, int]
  elements_per_block: int
  lds_bytes: int
  emitter: str
  schedule: dict = field(default_factory=dict)
  # HBM bytes per element the kernel is designed to move (compulsory traffic)
  bytes_per_element: int = 0
  shared_bytes: int = 0
  grid_override: str = ''       # C expression for the grid size, '' -> ceil(N / per block)
  extra_launch_setup: str = ''
  # a second kernel of the same chain, launched when ``condition`` (a C expression over the
  # launcher's parameters) is false
  fallback: 'Emitted' = None
  condition: str = ''

  def full_source(self):
    return self.kernel_src + (self.fallback.full_source() if self.fallback else '')


class Writer:
  """Indented source text builder."""

  def __init__(self, indent=2):
    self._lines: List[str] = []
    self._depth = 0
    self._indent = indent

  def __call__(self, line=''):
    if line:
      self._lines.append(' ' * (self._depth * self._indent) + line)
    else:
      self._lines.append('')

  def lines(self, block: str):
    for line in block.splitlines():
      self(line)

  class _Block:
    def __init__(self, writer, head, tail):
      self.w, self.head, self.tail = writer, head, tail

    def __enter__(self):
      self.w(self.head + ' {')
      self.w._depth += 1
      return self.w

    def __exit__(self, *exc):
      self.w._depth -= 1
      self.w('}' + self.tail)

  def block(self, head, tail=''):
    return Writer._Block(self, head, tail)

  def src(self):
    return '\n'.join(self._lines) + '\n'


def kernel_params(program, context):
  """Parameter list of the ``__global__`` kernel: the launcher's, minus ``streamPtr``."""
  params = []
  for a in program.args:
    dt = str(a.dtype)
    if a.addressing == Addressing.SCALAR:
      params.append(f'{dt} {a.name}')
    else:
      const = 'const ' if a.const else ''
      params.append(f'{const}{dt}{a.addressing.to_pointer()} __restrict__ {a.name}')
      if a.addressing != Addressing.NONE:
        params.append(f'unsigned {a.name}{GeneralLexicon.EXTRA_OFFSET}')
  params.append(f'size_t {GeneralLexicon.NUM_ELEMENTS}')
  params.append(f'const unsigned* __restrict__ {GeneralLexicon.FLAGS_NAME}')
  return params


def kernel_args(program):
  args = []
  for a in program.args:
    args.append(a.name)
    if a.addressing not in (Addressing.SCALAR, Addressing.NONE):
      args.append(f'{a.name}{GeneralLexicon.EXTRA_OFFSET}')
  args.append(GeneralLexicon.NUM_ELEMENTS)
  args.append(GeneralLexicon.FLAGS_NAME)
  return args


def launcher_source(generator, emitted: Emitted) -> str:
  w = Writer()
  proto = generator._generate_launcher_proto(with_defaults=False)
  with w.block(f'extern "C" {proto}'):
    w(f'if ({GeneralLexicon.NUM_ELEMENTS} == 0) return;')
    w(f'hipStream_t stream = ({GeneralLexicon.STREAM_PTR_STR} != nullptr) ? '
      f'static_cast<hipStream_t>({GeneralLexicon.STREAM_PTR_STR}) : 0;')
    if emitted.fallback is not None:
      with w.block(f'if ({emitted.condition})'):
        _launch_block(w, generator, emitted)
      with w.block('else'):
        _launch_block(w, generator, emitted.fallback)
    else:
      _launch_block(w, generator, emitted)
  return w.src()


def _launch_block(w, generator, emitted):
  program = generator.get_program()
  kname = emitted.kernel_name
  bx, by, bz = emitted.block
  w(f'dim3 block({bx}, {by}, {bz});')
  per_block = emitted.elements_per_block
  if emitted.extra_launch_setup:
    w.lines(emitted.extra_launch_setup)
  if emitted.grid_override:
    w(f'dim3 grid({emitted.grid_override}, 1, 1);')
  else:
    w(f'dim3 grid(static_cast<unsigned>(({GeneralLexicon.NUM_ELEMENTS} + {per_block} - 1) / '
      f'{per_block}), 1, 1);')
  if emitted.lds_bytes > 64 * 1024:
    with w.block('static bool shmemsizeset = false; if (!shmemsizeset)'):
      w(f'hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(&{kname}), '
        f'hipFuncAttributeMaxDynamicSharedMemorySize, {emitted.lds_bytes});')
      w(f'if (err != hipSuccess) {{ tensorforge_record_error(static_cast<int>(err), '
        f'"launcher_{generator.get_base_name()}"); return; }}')
      w('shmemsizeset = true;')
  args = ', '.join(kernel_args(program))
  w(f'hipLaunchKernelGGL({kname}, grid, block, {emitted.lds_bytes}, stream, {args});')
  w('hipError_t err = hipGetLastError();')
  w(f'if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), '
    f'"launcher_{generator.get_base_name()}");')


def translation_unit(generators) -> str:
  """One HIP source holding every generator's kernel and launcher. A name appears once; two
  different kernels under one name (descriptions that differ only where the name's hash does
  not look: floating type, sparsity pattern, constant values) cannot share an object."""
  from ..common.exceptions import GenerationError
  parts = [HEADER]
  seen = {}
  for g in generators:
    name = g.get_base_name()
    if name in seen:
      if seen[name] != (g.get_kernel(), g.get_launcher()):
        raise GenerationError(f'two different kernels are named {name}: build them into '
                              'separate objects')
      continue
    seen[name] = (g.get_kernel(), g.get_launcher())
    parts.append(g.get_kernel())
    parts.append(g.get_launcher())
  return '\n'.join(parts)
