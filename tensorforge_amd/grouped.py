"""One persistent launch over several generated chains.

SeisSol keeps the elements of each polynomial order (and of each time cluster) in their own
arrays and launches one generated kernel per chain; at low element counts per order the launches
are short and the GPU drains between them. ``GroupedGenerator`` emits one kernel that runs the
streaming schedule of every member chain in turn inside the same resident workgroups: each
workgroup stages member 0's shared operands in LDS, walks its share of member 0's element groups,
then moves on to member 1, and so on. No workgroup waits for another, so a workgroup that is done
with a short member starts on the next one while others still work.

The launcher is ``extern "C" void launcher_grouped_<md5>(<member 0 launcher parameters, names
prefixed g0_>, ..., <member n-1 parameters>, void* streamPtr)``: per member exactly the
parameters of its own launcher (``generators/generator.py:384-436`` of the reference), without
the stream, so a caller holding the member call sites passes the same arguments. This entry point
has no counterpart in the reference (its persistent option covers one chain,
``generators/generator.py:117-119``).
"""
import hashlib

from .common.basic_types import GeneralLexicon
from .common.exceptions import GenerationError
from .emit.launcher import HEADER, Writer, kernel_args, kernel_params, persistent_grid
from .emit.tuned import LDS_LIMIT


def _split_decl(decl):
  head, _, name = decl.rpartition(' ')
  return head, name


def _body(src, kname):
  """Lines of a stream kernel between its opening and closing brace, and the text before its
  signature (file-scope tables)."""
  lines = src.rstrip('\n').split('\n')
  sig = next(i for i, l in enumerate(lines) if l.startswith('extern "C" __global__') and
             f' {kname}(' in l)
  if lines[sig + 1].strip() != '{' or lines[-1].strip() != '}':
    raise GenerationError(f'unexpected layout of {kname}')
  return '\n'.join(lines[:sig]), lines[sig + 2:-1]


class GroupedGenerator:
  """Members: generated ``Generator`` objects whose tuned emitter chose the streaming schedule
  (same workgroup size)."""

  def __init__(self, generators):
    if not generators:
      raise ValueError('no member chains')
    self.members = list(generators)
    self._kernel = self._launcher = self._header = None
    self._name = None

  # ---------------------------------------------------------------- generation
  def generate(self):
    for g in self.members:
      if g.get_emitted() is None:
        g.generate()
    ems = [g.get_emitted() for g in self.members]
    for g, em in zip(self.members, ems):
      if em.schedule.get('kind') != 'stream' or em.fallback is not None:
        raise GenerationError(f'{g.get_base_name()}: grouped launches need the streaming '
                              f'schedule without alternative kernels (got {em.schedule})')
    threads = {em.block for em in ems}
    if len(threads) != 1:
      raise GenerationError(f'members use different workgroup sizes {sorted(threads)}')
    self.block = ems[0].block
    digest = hashlib.md5(', '.join(g.get_base_name() for g in self.members).encode()).hexdigest()
    self._name = f'grouped_{digest[:10]}'
    self.lds_bytes = max(em.lds_bytes for em in ems)
    if self.lds_bytes > LDS_LIMIT:
      raise GenerationError(f'{self.lds_bytes} bytes of LDS')
    self.per_block = [em.elements_per_block for em in ems]
    self._kernel = self._emit_kernel(ems)
    self._launcher = self._emit_launcher()
    self._header = f'{self._proto(with_defaults=False)};\n'

  def _member_params(self, i, kind):
    g = self.members[i]
    if kind == 'kernel':
      decls = kernel_params(g.get_program(), g._context)
    else:
      decls = [d.strip() for d in g._generate_base_params_list(g._global_symbols(),
                                                               with_types=True)]
    return [(h, n, f'g{i}_{n}') for h, n in (_split_decl(d) for d in decls)]

  def _emit_kernel(self, ems):
    w = Writer()
    kname = f'kernel_{self._name}'
    pre = []
    bodies = []
    for em in ems:
      p, b = _body(em.kernel_src, em.kernel_name)
      pre.append(p)
      bodies.append(b)
    for p in pre:
      if p.strip():
        w.lines(p)
    params = [f'{h} {pn}' for i in range(len(self.members))
              for h, _, pn in self._member_params(i, 'kernel')]
    mw = 1
    for em in ems:
      mw = max(mw, 2 if em.schedule.get('waves_per_block', 1) * 64 == self.block[0] else 1)
    w(f'extern "C" __global__ void __launch_bounds__({self.block[0]}, {mw}) '
      f'{kname}({", ".join(params)})')
    w('{')
    w._depth += 1
    for i, (g, body) in enumerate(zip(self.members, bodies)):
      w(f'// member {i}: {g.get_base_name()}')
      w('{')
      w._depth += 1
      for h, n, pn in self._member_params(i, 'kernel'):
        w(f'{h} {n} = {pn};')
      # a member without elements may pass pointers to empty arrays: its body (whose prefetch
      # reads element 0 before the element loop) must not run at all
      w(f'if ({GeneralLexicon.NUM_ELEMENTS} != 0)')
      w('{')
      w._depth += 1
      for line in body:
        w(line[2:] if line.startswith('  ') else line)
      w._depth -= 1
      w('}')
      w._depth -= 1
      w('}')
      if i + 1 < len(self.members):
        # the next member restages the shared operands in the same LDS
        w('__syncthreads();')
    w._depth -= 1
    w('}')
    self.kernel_name = kname
    return w.src()

  def _proto(self, with_defaults=False):
    params = [f'{h} {pn}' for i in range(len(self.members))
              for h, _, pn in self._member_params(i, 'launcher')]
    default = ' = nullptr' if with_defaults else ''
    params.append(f'void* {GeneralLexicon.STREAM_PTR_STR}{default}')
    return f'void launcher_{self._name}({", ".join(params)})'

  def _emit_launcher(self):
    w = Writer()
    kname = self.kernel_name
    threads = self.block[0]
    with w.block(f'extern "C" {self._proto()}'):
      w(f'hipStream_t stream = ({GeneralLexicon.STREAM_PTR_STR} != nullptr) ? '
        f'static_cast<hipStream_t>({GeneralLexicon.STREAM_PTR_STR}) : 0;')
      groups = ' + '.join(f'(g{i}_{GeneralLexicon.NUM_ELEMENTS} + {pb - 1}) / {pb}'
                          for i, pb in enumerate(self.per_block))
      w('if ((' + groups + ') == 0) return;')
      w.lines(persistent_grid(kname, self.lds_bytes, threads, 1, f'launcher_{self._name}',
                              groups_expr=groups))
      args = []
      for i, g in enumerate(self.members):
        args += [f'g{i}_{a}' for a in kernel_args(g.get_program())]
      w(f'hipLaunchKernelGGL({kname}, dim3(tf_grid, 1, 1), dim3({threads}, 1, 1), '
        f'{self.lds_bytes}, stream, {", ".join(args)});')
      w('hipError_t err = hipGetLastError();')
      w(f'if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), '
        f'"launcher_{self._name}");')
    return w.src()

  # ---------------------------------------------------------------- access (Generator-like)
  def get_base_name(self):
    return self._name

  def get_kernel(self):
    return self._kernel

  def get_launcher(self):
    return self._launcher

  def get_header(self):
    return self._header

  def get_source(self):
    return HEADER + self._kernel + self._launcher

  def launcher_signature(self):
    out = []
    for i, g in enumerate(self.members):
      out += [(t, f'g{i}_{n}') for t, n in g.launcher_signature()[:-1]]
    out.append(('void*', GeneralLexicon.STREAM_PTR_STR))
    return out

  def member_args(self, member_args):
    """Grouped launcher arguments from each member's own launcher argument list (stream last,
    as ``device.launch_args`` builds it); the stream of the first member is used."""
    if len(member_args) != len(self.members):
      raise ValueError(f'{len(self.members)} members, {len(member_args)} argument lists')
    out = []
    for a in member_args:
      out += list(a[:-1])
    out.append(member_args[0][-1])
    return out
