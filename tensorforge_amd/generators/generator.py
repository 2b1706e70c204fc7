"""``Generator``: one fused gfx950 kernel and its C launcher from a description list.

Public surface as in the reference (``tensorforge/generators/generator.py:52-492``): operand
naming, kernel-name hashing, launcher prototype, header and call-site strings are restated so
that generated call sites stay valid. What ``generate()`` emits is new: gfx950 HIP source from one
of two emitters (:mod:`tensorforge_amd.emit.plain` restating the reference's thread mapping, or
:mod:`tensorforge_amd.emit.tuned`), with an ``extern "C"`` launcher.
"""
import hashlib
from copy import deepcopy
from typing import List, Type, Union

from ..common.basic_types import Addressing, DataFlowDirection, GeneralLexicon
from ..common.context import Context
from ..common.exceptions import GenerationError
from ..emit import ir as ir_mod
from .descriptions import ElementwiseDescr, MultilinearDescr, OperationDescription


class AbstractThreadBlockPolicy:
  def __init__(self, context: Context, global_mem: int, mem_per_mult: int, num_threads: int):
    self._context = context
    self._mem_per_mult = mem_per_mult
    self._global_mem = global_mem
    self._num_threads = num_threads
    hw = context.get_vm().get_hw_descr()
    self._max_blocks = hw.max_block_per_sm
    self._max_allowed_mem = hw.max_local_mem_size_per_block
    self._max_threads = hw.max_threads_per_block

  def get_num_mults_per_block(self):
    raise NotImplementedError


class RegmaxBlockPolicy(AbstractThreadBlockPolicy):
  """Elements per block: ``256 // lanes``, capped by LDS (reference ``generator.py:38-50``)."""

  def get_num_mults_per_block(self):
    max_thread_mults = max(1, 256 // self._num_threads)
    if self._mem_per_mult == 0:
      return max_thread_mults
    size = self._context.fp_type.size()
    max_mem_mults = (self._max_allowed_mem - self._global_mem * size) // (self._mem_per_mult * size)
    return max(1, min(max_mem_mults, max_thread_mults))


class _GlobalSymbol:
  """What ``default_generate_call_site`` deep-copies and renames; mirrors the reference's symbol
  fields that the call-site helpers read (``name`` and ``obj``)."""

  def __init__(self, tensor):
    self.name = tensor.name
    self.obj = tensor


class Generator:
  NAME_ENCODING_LENGTH = 10

  def __init__(self, gemm_list: List[OperationDescription], context: Context,
               thread_block_policy_type: Type[AbstractThreadBlockPolicy] = RegmaxBlockPolicy,
               emitter: str = 'auto', schedule=None):
    """``emitter``: ``'plain'`` (reference thread mapping), ``'tuned'`` (gfx950 schedule; raises
    if the chain is outside its scope) or ``'auto'`` (tuned when applicable). ``schedule`` is an
    optional :class:`tensorforge_amd.emit.tuned.Schedule` overriding the default choice."""
    self.descr_list = gemm_list
    self._context = context
    self._thread_block_policy_type = thread_block_policy_type
    self._base_kernel_name: Union[str, None] = None
    self._emitter_choice = emitter
    self._schedule = schedule
    self._kernel = None
    self._launcher = None
    self._header = None
    self._matrix_list = None
    self._tmp_list = None
    self._global_tensors = []
    self._is_registerd = False
    self._num_threads = 0
    self._num_active_threads = 0
    self._program = None
    self._emitted = None
    self._persistent_threading = False
    self._preload_globals = False
    self._check_consistency_with_user_options()
    self._name_operands(self.descr_list)

  # ---- naming and registration (reference generator.py:88-96, 292-362) ----
  def set_kernel_name(self, name):
    self._base_kernel_name = name

  def register(self):
    self._collect_tmp_matrices()
    self._populate_global_scope()
    if not self._base_kernel_name:
      self._generate_kernel_name()
    self._is_registerd = True

  def _check_consistency_with_user_options(self):
    user_options = self._context.get_user_options()
    for descr in self.descr_list:
      if not descr.is_strict_match() == user_options.exact_contraction_length:
        raise RuntimeError('gemm list is not consistent with user options. '
                           f'`strict_math` in gemm descr. set to {descr.is_strict_match()}, '
                           'but `exact_contraction_length` is set to '
                           f'{user_options.exact_contraction_length}')

  def _name_operands(self, gemm_list):
    ordered = {}
    for descr in gemm_list:
      for sub in descr.matrix_list():
        ordered[sub.tensor] = None
    self._matrix_list = list(ordered.keys())
    tmp_counter = op_counter = 0
    for tensor in self._matrix_list:
      if tensor.is_tmp:
        tensor.name = f't{tmp_counter}'
        tmp_counter += 1
      else:
        tensor.name = f'm{op_counter}'
        op_counter += 1

  def _collect_tmp_matrices(self):
    self._tmp_list = []
    for tensor in self._matrix_list:
      if tensor.is_tmp and tensor not in self._tmp_list:
        self._tmp_list.append(tensor)

  def _populate_global_scope(self):
    self._global_tensors = [t for t in self._matrix_list if t not in self._tmp_list]

  def _generate_kernel_name(self):
    long_name = [t.gen_descr() for t in self._global_tensors]
    long_name.extend(str(descr) for descr in self.descr_list)
    digest = hashlib.md5(', '.join(long_name).encode()).hexdigest()
    self._base_kernel_name = f'kernel_{digest[:Generator.NAME_ENCODING_LENGTH]}'

  def get_base_name(self):
    return self._base_kernel_name

  # ---- generation ----
  def generate(self):
    if not self._is_registerd:
      self.register()
    pointwise = [isinstance(d, ElementwiseDescr) for d in self.descr_list]
    tail = pointwise.index(True) if any(pointwise) else len(pointwise)
    fused = 0 < tail < len(pointwise) and all(pointwise[tail:])
    if any(pointwise) and not all(pointwise) and not fused:
      raise GenerationError('pointwise statements (ElementwiseDescr) come after the contraction '
                            'chain they fuse into, or stand alone')
    self._deduce_num_threads()
    from ..emit import plain, tuned, launcher
    if any(pointwise) and self._context.backend == 'omptarget':
      raise GenerationError('the omptarget backend generates multilinear chains only')
    if pointwise and all(pointwise):
      from ..emit import pointwise as pw
      if self._emitter_choice not in ('auto', 'tuned', 'plain'):
        raise GenerationError(f'{self._emitter_choice} does not generate pointwise kernels')
      self._program = pw.lower(self.descr_list, self._global_tensors, self._context.fp_type)
      emitted = launcher.distinct_symbol(pw.emit(self._program, self._context,
                                                 self._base_kernel_name))
      self._emitted = emitted
      self._kernel = emitted.full_source()
      self._launcher = launcher.launcher_source(self, emitted)
      self._header = f'{self._generate_launcher_proto(with_defaults=True)};\n'
      return
    self._program = ir_mod.lower(self.descr_list[:tail], self._global_tensors,
                                 self._context.fp_type)
    if self._context.backend == 'omptarget':
      # the CPU form: one host C++/OpenMP launcher with the same name and parameters
      from ..emit import omp
      if fused:
        raise GenerationError('the omptarget backend generates multilinear chains only')
      self._emitted = launcher.Emitted(kernel_name=f'launcher_{self._base_kernel_name}',
                                       kernel_src='', block=(1, 1, 1), elements_per_block=1,
                                       lds_bytes=0, emitter='omptarget')
      self._kernel = ''
      self._launcher = omp.emit(self._program, self)
      self._header = f'{self._generate_launcher_proto(with_defaults=True)};\n'
      return
    if fused:
      # the statements run on the chain's destination in registers, before it is stored
      from ..emit import pointwise as pw
      self._program.epilogue = pw.lower(self.descr_list[tail:], self._global_tensors,
                                        self._context.fp_type)
      if self._emitter_choice not in ('auto', 'tuned'):
        raise GenerationError(f'{self._emitter_choice} does not fuse pointwise statements')
    choice = self._emitter_choice
    emitted = None
    if choice in ('mfma_fp64', 'mfma_fp32', 'mfma_bf16x3'):
      from ..emit import mfma
      emitted, reason = mfma.emit(self._program, self._context, self._base_kernel_name,
                                  choice[len('mfma_'):], self._schedule)
      if emitted is None:
        raise GenerationError(f'{choice} cannot generate this chain: {reason}')
    if choice in ('auto', 'tuned'):
      reason = tuned.unsupported_reason(self._program, self._context, self._schedule)
      if reason is None:
        try:
          emitted = tuned.emit(self._program, self._context, self._base_kernel_name,
                               self._schedule)
        except GenerationError:
          if choice == 'tuned':
            raise
        if choice == 'auto' and emitted is not None and tuned.starved(emitted) and \
            self._program.epilogue is None:
          emitted = None      # the plain kernel keeps more wavefronts resident
      elif choice == 'tuned':
        raise GenerationError(f'tuned emitter cannot generate this chain: {reason}')
    if emitted is None:
      if choice not in ('auto', 'tuned', 'plain', 'mfma_fp64', 'mfma_fp32', 'mfma_bf16x3'):
        raise ValueError(f'unknown emitter {choice!r}')
      emitted = plain.emit(self._program, self._context, self._base_kernel_name,
                           self._num_threads, self._thread_block_policy_type)
    emitted = launcher.distinct_symbol(emitted)
    self._emitted = emitted
    self._kernel = emitted.full_source()
    self._launcher = launcher.launcher_source(self, emitted)
    self._header = f'{self._generate_launcher_proto(with_defaults=True)};\n'

  def _deduce_num_threads(self):
    for descr in self.descr_list:
      num_threads, num_active = descr.get_num_threads(self._context)
      self._num_threads = max(num_threads, self._num_threads)
      self._num_active_threads = max(num_active, self._num_active_threads)
    if not any(isinstance(d, ElementwiseDescr) for d in self.descr_list):
      self._num_threads = min(32, self._num_threads)

  def get_kernel(self):
    return self._kernel

  def get_launcher(self):
    return self._launcher

  def get_header(self):
    return self._header

  def get_program(self):
    return self._program

  def get_emitted(self):
    """Emitter output: schedule, launch geometry and per-element traffic figures."""
    return self._emitted

  def get_source(self):
    """A complete translation unit: includes, kernel, ``extern "C"`` launcher (HIP; host
    C++/OpenMP for the omptarget backend)."""
    if self._context.backend == 'omptarget':
      return self._launcher
    from ..emit import launcher
    return launcher.translation_unit([self])

  def get_helper_headers(self):
    return []

  # ---- prototypes and call sites (reference generator.py:384-492) ----
  def _arg_type(self, tensor):
    return self._context.fp_type if tensor.datatype is None else tensor.datatype

  def _generate_base_params_list(self, symbol_list, with_types=True, with_defaults=False):
    params = []
    for symbol in symbol_list:
      tensor = symbol.obj
      datatype = self._arg_type(tensor)
      if tensor.addressing == Addressing.SCALAR:
        params.append(f'{datatype} {symbol.name}' if with_types else f'{symbol.name}')
      else:
        const = 'const ' if tensor.direction == DataFlowDirection.SOURCE else ''
        batch_type = f'{const}{datatype}{tensor.addressing.to_pointer()}' if with_types else ''
        offset_type = 'unsigned' if with_types else ''
        params.append(f'{batch_type} {symbol.name}')
        if tensor.addressing != Addressing.NONE:
          params.append(f'{offset_type} {symbol.name}{GeneralLexicon.EXTRA_OFFSET}')
    params.append(f'{"size_t" if with_types else ""} {GeneralLexicon.NUM_ELEMENTS}')
    flags_type = 'unsigned*' if with_types else ''
    default_flags = '= nullptr' if with_defaults else ''
    params.append(f'{flags_type} {GeneralLexicon.FLAGS_NAME} {default_flags}')
    return params

  def _global_symbols(self):
    return [_GlobalSymbol(t) for t in self._global_tensors]

  def _generate_launcher_proto(self, with_defaults=True):
    params = self._generate_base_params_list(self._global_symbols(), with_types=True,
                                             with_defaults=with_defaults)
    default_value = ' = nullptr' if with_defaults else ''
    params.append(f'void* {GeneralLexicon.STREAM_PTR_STR}{default_value}')
    return f'void launcher_{self._base_kernel_name}({", ".join(params)})'

  def default_generate_call_site(self):
    """The reference's default call-site string (``generator.py:445-459``): global operands
    under their aliases, then ``flags, streamPtr`` appended to a parameter list that already
    ends in ``numElements, flags`` (so ``flags`` appears twice, as there)."""
    if not self._is_registerd:
      raise RuntimeError('generator is not registered. Call register first.')
    symbols = deepcopy(self._global_symbols())
    for item in symbols:
      if item.obj.alias:
        item.name = item.obj.alias
    args = self._generate_base_params_list(symbols, with_types=False)
    args.append(GeneralLexicon.FLAGS_NAME)
    args.append(GeneralLexicon.STREAM_PTR_STR)
    return f'launcher_{self._base_kernel_name}({", ".join(args)});'

  def generate_call_site(self, mat_name_map, offset_name_map, num_element, flags=None,
                         stream=None):
    args = []
    for tensor in self._global_tensors:
      if tensor.alias in mat_name_map:
        args.append(mat_name_map[tensor.alias])
        if tensor.addressing not in (Addressing.SCALAR, Addressing.NONE):
          args.append(offset_name_map[tensor.alias])
    args.append(num_element)
    if flags:
      args.append(flags)
    if stream:
      args.append(stream)
    return f'launcher_{self._base_kernel_name}({", ".join(args)});'

  def launcher_signature(self):
    """``[(c_type, name)]`` of the launcher, in order (used by the ctypes loader)."""
    out = []
    for tensor in self._global_tensors:
      dt = str(self._arg_type(tensor))
      if tensor.addressing == Addressing.SCALAR:
        out.append((dt, tensor.name))
      else:
        out.append((f'{dt}{tensor.addressing.to_pointer()}', tensor.name))
        if tensor.addressing != Addressing.NONE:
          out.append(('unsigned', f'{tensor.name}{GeneralLexicon.EXTRA_OFFSET}'))
    out.append(('size_t', GeneralLexicon.NUM_ELEMENTS))
    out.append(('unsigned*', GeneralLexicon.FLAGS_NAME))
    out.append(('void*', GeneralLexicon.STREAM_PTR_STR))
    return out
