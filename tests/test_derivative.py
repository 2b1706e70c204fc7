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


def test_three_streamed_arrays_fall_back_to_the_plain_kernel():
  """Three streamed arrays are outside the streaming form; the column-per-lane form needs
  about 82 KB of LDS per one-wavefront workgroup (one wavefront per CU), so the default choice
  is the plain kernel; asking for the tuned emitter still gives the column-per-lane form."""
  from tensorforge_amd.emit import tuned
  w = W.ader_derivative(aligned=False)
  ctx = Context(arch='gfx950', backend='hip', fp_type=FloatingPointType.DOUBLE)
  g = Generator(w.descrs, ctx)
  g.generate()
  assert g.get_emitted().emitter == 'plain'
  t = Generator(w.descrs, ctx, emitter='tuned')
  t.generate()
  em = t.get_emitted()
  assert em.emitter == 'tuned' and em.schedule.get('kind') != 'stream'
  assert tuned.starved(em)
  assert not tuned.starved(_emitted(W.ader_derivative(aligned=True), ctx))


def _emitted(workload, ctx):
  g = Generator(workload.descrs, ctx)
  g.generate()
  return g.get_emitted()
