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
      # the failed call leaves HIP's last error set: clear it, or the next launcher that
      # succeeds would read it back as its own failure
      w(f'if (err != hipSuccess) {{ (void)hipGetLastError(); '
        f'tensorforge_record_error(static_cast<int>(err), '
        f'"launcher_{generator.get_base_name()}"); return; }}')
      w('shmemsizeset = true;')
  args = ', '.join(kernel_args(program))
  w(f'hipLaunchKernelGGL({kname}, grid, block, {emitted.lds_bytes}, stream, {args});')
  w('hipError_t err = hipGetLastError();')
  w(f'if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), '
    f'"launcher_{generator.get_base_name()}");')


def persistent_grid(kname: str, lds_bytes: int, threads: int, per_block: int, where: str,
                    groups_expr: str = '') -> str:
  """C statements sizing the grid of a persistent kernel: as many workgroups as stay resident
  (CUs x workgroups per CU, queried once), at most one per ``per_block`` elements. Every HIP
  call's status is checked: on a failure (or a kernel that cannot be resident at all) the error
  is recorded, HIP's last error is cleared so later launchers do not read it back, and the
  launcher returns without launching. ``groups_expr`` overrides the number of workgroup-sized
  element groups."""
  return (f'static int tf_resident = 0;\n'
          f'if (tf_resident == 0) {{\n'
          f'  int dev = 0, cus = 0, per_cu = 0;\n'
          f'  hipError_t qerr = hipGetDevice(&dev);\n'
          f'  if (qerr == hipSuccess) qerr = hipDeviceGetAttribute(&cus, '
          f'hipDeviceAttributeMultiprocessorCount, dev);\n'
          f'  if (qerr == hipSuccess) qerr = hipFuncSetAttribute(reinterpret_cast<const void*>('
          f'&{kname}), hipFuncAttributeMaxDynamicSharedMemorySize, {lds_bytes});\n'
          f'  if (qerr == hipSuccess) qerr = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, '
          f'{kname}, {threads}, {lds_bytes});\n'
          f'  if (qerr == hipSuccess && (cus <= 0 || per_cu <= 0)) qerr = hipErrorInvalidConfiguration;\n'
          f'  if (qerr != hipSuccess) {{\n'
          f'    (void)hipGetLastError();\n'
          f'    tensorforge_record_error(static_cast<int>(qerr), "{where}");\n'
          f'    return;\n'
          f'  }}\n'
          f'  tf_resident = cus * per_cu;\n'
          f'}}\n'
          + (f'const size_t tf_groups = {groups_expr};\n' if groups_expr else
             f'const size_t tf_groups = (numElements + {per_block - 1}) / {per_block};\n') +
          f'const unsigned tf_grid = static_cast<unsigned>(tf_groups < (size_t)tf_resident ? '
          f'tf_groups : (size_t)tf_resident);')


def distinct_symbol(emitted: Emitted) -> Emitted:
  """Rename the ``__global__`` symbol of a non-plain kernel to
  ``<name>_<form>_<digest of its source>``. The launcher keeps the reference's name; the plain
  kernel keeps the reference's kernel name. Different forms (and schedules) of one chain then
  export different kernel symbols, so two objects holding forms of one chain can be loaded into
  one process without their kernels or their launchers' one-time statics being confused, and a
  profiler's trace names each form."""
  import dataclasses
  import hashlib
  if emitted.emitter == 'plain':
    return emitted
  form = str(emitted.schedule.get('kind', emitted.emitter))
  digest = hashlib.sha1(emitted.full_source().encode()).hexdigest()[:8]
  old = emitted.kernel_name
  new = f'{old}_{form}_{digest}'

  def rename(em):
    fb = rename(em.fallback) if em.fallback is not None else None
    return dataclasses.replace(em, kernel_name=em.kernel_name.replace(old, new),
                               kernel_src=em.kernel_src.replace(old, new),
                               extra_launch_setup=em.extra_launch_setup.replace(old, new),
                               condition=em.condition.replace(old, new), fallback=fb)
  return rename(emitted)


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
