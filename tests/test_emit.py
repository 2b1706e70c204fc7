"""Structure of the generated kernels (CPU: generation only)."""
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
  src = g.get_kernel()
  assert src.count('v_fmac_f64_dpp') == nnz
  assert em.schedule['shared_entries'] == nnz
  # no scalar-memory reads of the shared matrices inside the loop: they come from LDS chunks
  assert 'ubuf[' in src


def test_stream_schedule_recognises_the_single_gemm():
  g = _gen(W.gemm())
  assert g.get_emitted().schedule['kind'] == 'stream'
  assert g.get_kernel().count('v_fmac_f64_dpp') == 56 * 56


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
