"""The time-derivative step (reference ``benchmarks/fused/programs/seissol-deriv_align.cf`` and
``seissol-deriv_no_align.cf``) on the CPU: the oracle against a closed form written from the
programs, and the streaming plan for terms that contract over different leading row ranges of
one streamed array."""
import numpy as np
import pytest

from tensorforge_amd import catalog, workloads as W
from tensorforge_amd.common.basic_types import FloatingPointType
from tensorforge_amd.common.context import Context
from tensorforge_amd.common.exceptions import GenerationError
from tensorforge_amd.generators.generator import Generator

from harness import make_inputs, oracle_outputs


def _closed_form(host, n, aligned):
  """``D[:rows] += K_d[:rows, :k_d] · (A_d[window] · S_d)``; a window's first row sits at the
  array's first storage row (pointers address the bounding box's lower corner)."""
  rows = 48 if aligned else 35
  D = host['D'].reshape(n, 9, 48).transpose(0, 2, 1).copy()
  for d, kc in enumerate((53, 54, 55)):
    A = (host['A'] if aligned else host[f'A{d + 1}']).reshape(n, 9, 64).transpose(0, 2, 1)
    S = host[f'S{d + 1}'].reshape(n, 9, 9).transpose(0, 2, 1)
    K = host[f'K{d + 1}'].reshape(kc, 48).T
    D[:, :rows] += np.einsum('ik,ekj->eij', K[:rows, :kc], A[:, :kc] @ S)
  return D


@pytest.mark.parametrize('aligned', [True, False])
def test_oracle_matches_the_closed_form(aligned):
  entry = catalog.get('deriv_f64_plain' if aligned else 'deriv_na_f64_plain')
  entry.generator.generate()
  n = 4
  host = make_inputs(entry, n, seed=3)
  got = oracle_outputs(entry, host, n)['D'].reshape(n, 9, 48).transpose(0, 2, 1)
  want = _closed_form(host, n, aligned)
  assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
  if not aligned:
    before = host['D'].reshape(n, 9, 48).transpose(0, 2, 1)
    assert np.array_equal(got[:, 35:], before[:, 35:])


def test_streaming_plan_keeps_each_term_in_its_range():
  from tensorforge_amd.emit import stream
  w = W.ader_derivative(aligned=True)
  g = Generator(w.descrs, Context(arch='gfx950', backend='hip',
                                  fp_type=FloatingPointType.DOUBLE))
  g.generate()
  em = g.get_emitted()
  assert em.emitter == 'tuned' and em.schedule['kind'] == 'stream'
  assert em.schedule['shared_entries'] == 48 * (53 + 54 + 55)
  pl, why = stream.plan(g._program)
  assert why is None and (pl.klo, pl.khi) == (0, 55)
  # the unrolled form: a runtime block loop would stage every term over the widest range
  assert not em.schedule.get('loop')
  # matrix-core forms read every term over the plan's range: refused
  with pytest.raises(GenerationError, match='different contraction ranges'):
    Generator(w.descrs, Context(arch='gfx950', backend='hip', fp_type=FloatingPointType.DOUBLE),
              emitter='mfma_fp32').generate()


def test_three_streamed_arrays_stream_one_term_after_the_other():
  """Three streamed arrays (the step that is not aligned): the streaming form runs each term's
  contraction blocks in turn, loading each block's rows from its own array (row 1 of the array
  is contraction index 0), so the default choice is the streaming kernel. Without the streaming
  form the column-per-lane form needs about 82 KB of LDS per one-wavefront workgroup (one
  wavefront per CU) and counts as starved."""
  from tensorforge_amd.emit import stream, tuned
  from tensorforge_amd.emit.tuned import Schedule
  w = W.ader_derivative(aligned=False)
  ctx = Context(arch='gfx950', backend='hip', fp_type=FloatingPointType.DOUBLE)
  g = Generator(w.descrs, ctx)
  g.generate()
  em = g.get_emitted()
  assert em.emitter == 'tuned' and em.schedule.get('kind') == 'stream'
  pl, _ = stream.plan(g.get_program())
  assert pl.stacked and [t.l_shift for t in pl.terms] == [1, 1, 1]
  assert [t.L.arg.name for t in pl.terms] == sorted({t.L.arg.name for t in pl.terms},
                                                    key=[t.L.arg.name for t in pl.terms].index)
  src = g.get_kernel()
  # blocks of term 0, then 1, then 2; each term's first block read from its own array
  assert src.index('of term 0') < src.index('of term 1') < src.index('of term 2')
  assert 'lp_cur1' in src and 'lp_cur2' in src
  c = Generator(w.descrs, ctx, emitter='tuned', schedule=Schedule(stream=False))
  c.generate()
  em = c.get_emitted()
  assert em.emitter == 'tuned' and em.schedule.get('kind') != 'stream'
  assert tuned.starved(em)
  assert not tuned.starved(_emitted(W.ader_derivative(aligned=True), ctx))


def _emitted(workload, ctx):
  g = Generator(workload.descrs, ctx)
  g.generate()
  return g.get_emitted()
