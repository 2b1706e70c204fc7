"""Structure of the generated kernels (CPU: generation only)."""
import re
from tensorforge_amd import workloads as W
from tensorforge_amd.common.basic_types import FloatingPointType
from tensorforge_amd.common.context import Context
from tensorforge_amd.emit import stream
from tensorforge_amd.emit.tuned import Schedule
from tensorforge_amd.generators.generator import Generator


def _gen(w, fp=FloatingPointType.DOUBLE, **kw):
  g = Generator(w.descrs, Context('gfx950', 'hip', fp), emitter='tuned', schedule=Schedule(**kw))
  g.generate()
  return g


def test_stream_schedule_one_fma_per_stored_entry():
  w = W.ader_volume()
  g = _gen(w)
  em = g.get_emitted()
  assert em.schedule['kind'] == 'stream'
  nnz = sum(int(m.sum()) for m in w.meta['masks'])
  kernels = [em] + ([em.fallback] if em.fallback else [])
  for k in kernels:
    assert k.kernel_src.count('v_fmac_f64_dpp') == nnz
  assert em.schedule['shared_entries'] == nnz
  src = g.get_kernel()
  # no scalar-memory reads of the shared matrices inside the loop: they come from LDS chunks
  assert 'ubuf[' in src


def test_stream_schedule_recognises_the_single_gemm():
  g = _gen(W.gemm())
  assert g.get_emitted().schedule['kind'] == 'stream'
  assert g.get_emitted().kernel_src.count('v_fmac_f64_dpp') == 56 * 56


def test_stream_plan_rejects_other_shapes():
  from tensorforge_amd import catalog
  e = catalog.get('t_five_multiplies_auto')
  e.generator.register()
  from tensorforge_amd.emit import ir
  prog = ir.lower(e.generator.descr_list, e.generator._global_tensors, e.fp_type)
  pl, why = stream.plan(prog)
  assert pl is None and why


def test_fp32_stream_uses_f32_dpp():
  g = _gen(W.ader_volume(FloatingPointType.FLOAT), FloatingPointType.FLOAT)
  assert 'v_fmac_f32_dpp' in g.get_kernel()


def test_plain_schedule_follows_reference_mapping():
  g = Generator(W.ader_volume().descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE),
                emitter='plain')
  g.generate()
  em = g.get_emitted()
  assert em.block == (32, 8, 1) and em.emitter == 'plain'


def test_stream_direct_lds_ring_has_an_unaligned_fallback():
  """The direct memory-to-LDS ring needs 16-byte aligned rows: the launcher checks the streamed
  operand's base (pointer + extraOffset) and launches the register-ring kernel otherwise."""
  g = _gen(W.ader_volume(), stream_dma=True)
  em = g.get_emitted()
  assert em.schedule.get('lds_dma_slots', 0) >= 2
  assert 'global_load_lds' in em.kernel_src and 'global_load_lds' not in em.fallback.kernel_src
  assert em.fallback.kernel_name == em.kernel_name + '_r'
  launcher = g.get_launcher()
  assert '% 16u == 0u' in launcher and em.fallback.kernel_name in launcher
  # odd row counts cannot use 16-byte row pairs: no direct ring
  g5 = _gen(W.ader_volume(b=35, sparse_k=False), stream_dma=True, stream_loop='never')
  assert 'lds_dma_slots' not in g5.get_emitted().schedule


def test_flux_chains_stream_with_register_stages():
  """D += U2 · (U1 · (P · A)): the innermost product streams, the outer ones run on registers;
  every stored entry of each shared matrix's window is one DPP FMA."""
  for wl in (W.flux_local(), W.flux_neigh()):
    g = _gen(wl)
    em = g.get_emitted()
    assert em.schedule['kind'] == 'stream'
    pl, _ = stream.plan(g.get_program())
    assert pl is not None and len(pl.post) >= 1
    windows = {'local': [(21, 56), (56, 21)], 'neigh': [(21, 56), (21, 21), (56, 21)]}
    kind = 'local' if 'i_surf' in wl.tensors else 'neigh'
    expected = sum(r * c for r, c in windows[kind])
    assert em.kernel_src.count('v_fmac_f64_dpp') == expected


def test_row_bands_for_tall_dense_destinations():
  g = _gen(W.ader_volume(b=84, sparse_k=False))
  em = g.get_emitted()
  assert em.schedule.get('row_passes') == 2 and em.schedule.get('loop')
  assert em.lds_bytes <= 160 * 1024


def test_row_bands_of_an_unrolled_kernel_have_their_own_tables():
  """Each band stages its own rows of the shared matrices: a shared table name would make the
  second band read the first band's entries."""
  g = _gen(W.ader_volume(b=84, sparse_k=False), stream_loop='never', stream_pass_rows=48)
  src = g.get_emitted().kernel_src
  assert g.get_emitted().schedule.get('row_passes') == 2
  names = set(re.findall(r'static __device__ const unsigned (\w+)\[', src))
  assert len(names) == 2
  for n in names:
    assert f'{n}[e]' in src


def test_matrix_core_forms_refuse_register_stage_chains():
  """The flux integrals stream an inner product and run outer stages on registers, with a final
  accumulation into the destination: the matrix-core forms do not implement those stages and
  must refuse the chains instead of generating a kernel that drops them."""
  import pytest
  from tensorforge_amd.common.exceptions import GenerationError
  for make in (W.flux_local, W.flux_neigh):
    w = make()
    for mode in ('mfma_fp64', 'mfma_fp32', 'mfma_bf16x3'):
      g = Generator(w.descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE), emitter=mode)
      with pytest.raises(GenerationError, match='register stages'):
        g.generate()


def test_fp64_matrix_core_form_shapes():
  """The FP64 matrix-core form: the GEMM runs four wavefronts with 16-row blocks and the
  destination's old values in registers of their own, the dense order-6 chain four wavefronts
  with 8-row blocks (16 would not fit in LDS as doubles); schedule knobs override the shape and
  the destination mode; order 7 does not fit, and sparse shared operands, FP32 chains and shapes
  outside the form are refused."""
  import pytest
  from tensorforge_amd.common.exceptions import GenerationError
  ctx = Context('gfx950', 'hip', FloatingPointType.DOUBLE)
  g = Generator(W.gemm().descrs, ctx, emitter='mfma_fp64')
  g.generate()
  sch = g.get_emitted().schedule
  assert (sch['kind'], sch['waves_per_block'], sch['k_block']) == ('mfma_fp64', 4, 16)
  assert sch['lds_bytes'] <= 160 * 1024
  assert '__builtin_amdgcn_mfma_f64_16x16x4f64' in g.get_kernel()
  assert 'v += cpre[' in g.get_kernel()
  from tensorforge_amd.emit.tuned import Schedule
  g = Generator(W.gemm().descrs, ctx, emitter='mfma_fp64',
                schedule=Schedule(mfma_waves=8, mfma_k_block=8, mfma_dest='acc'))
  g.generate()
  sch = g.get_emitted().schedule
  assert (sch['waves_per_block'], sch['k_block']) == (8, 8) and 'cpre' not in g.get_kernel()
  for bad in (dict(mfma_waves=2, mfma_k_block=8), dict(mfma_dest='later')):
    g = Generator(W.gemm().descrs, ctx, emitter='mfma_fp64', schedule=Schedule(**bad))
    with pytest.raises(GenerationError):
      g.generate()
  g = Generator(W.ader_volume(sparse_k=False).descrs, ctx, emitter='mfma_fp64')
  g.generate()
  sch = g.get_emitted().schedule
  assert (sch['waves_per_block'], sch['k_block']) == (4, 8) and sch['lds_bytes'] <= 160 * 1024
  for w, why in ((W.ader_volume(b=84, sparse_k=False), 'do not fit in LDS'),
                 (W.ader_volume(), 'sparse shared operand')):
    g = Generator(w.descrs, ctx, emitter='mfma_fp64')
    with pytest.raises(GenerationError, match=why):
      g.generate()
  g = Generator(W.gemm(FloatingPointType.FLOAT).descrs,
                Context('gfx950', 'hip', FloatingPointType.FLOAT), emitter='mfma_fp64')
  with pytest.raises(GenerationError, match='FP64 chains'):
    g.generate()


def test_forms_of_one_chain_export_distinct_kernel_symbols():
  """The launcher keeps the reference's name for every form; the plain kernel keeps the
  reference's kernel name; other forms get a kernel symbol of their own (form and source
  digest), so forms of one chain can share a process and a trace."""
  w = W.ader_volume(b=84, sparse_k=False, star=True)
  names = {}
  for emitter in ('plain', 'tuned', 'mfma_fp32', 'mfma_bf16x3'):
    g = Generator(w.descrs, Context('gfx950', 'hip', FloatingPointType.DOUBLE), emitter=emitter)
    g.generate()
    em = g.get_emitted()
    names[emitter] = em.kernel_name
    assert f'launcher_{g.get_base_name()}(' in g.get_launcher()
    assert re.search(rf'__global__ void __launch_bounds__\([^)]*\) {em.kernel_name}\(',
                     g.get_kernel())
    assert f'&{em.kernel_name})' in g.get_launcher() or f'({em.kernel_name},' in g.get_launcher()
  base = 'kernel_' + g.get_base_name()
  assert names['plain'] == base
  assert names['tuned'].startswith(base + '_stream_')
  assert names['mfma_fp32'].startswith(base + '_mfma_fp32_')
  assert names['mfma_bf16x3'].startswith(base + '_mfma_bf16x3_')
  assert len(set(names.values())) == 4
