"""``tune.autotune``: the cost model's pick, the pruned search and the exhaustive sweep on chains
outside the timing database (three random chains of the cost model's workload, seed 2026
indices 40000-40002, and the order-6 volume chain at 262,144 elements with sparse and with dense
K_d), candidates prebuilt by ``tests/autotune_build.py``.

Each method returns a candidate the chain admits, with a generated launcher; the pruned search
times at most three candidates and keeps the fastest of them; the exhaustive sweep times all of
them. On the volume chain the pick and the pruned search land within 5 % of the sweep's best.
The FP64 matrix-core form, which the model was not trained on, is timed by both timed methods
on the dense chain and never picked untimed. The slowdowns of every chain are printed.
"""
import pytest

import autotune_build as ab

pytestmark = pytest.mark.gpu


def test_autotune_methods(device):
  from tensorforge_amd.tune.autotune import autotune
  for name, descrs, fp, batch in ab.chains():
    ex = autotune(descrs, fp, batch, method='exhaustive', build_dir=ab.BUILD_DIR)
    pick = autotune(descrs, fp, batch, method='pick')
    pr = autotune(descrs, fp, batch, method='pruned', k=3, build_dir=ab.BUILD_DIR)
    best = min(ex.timed.values())
    extra = set(ex.timed) - set(ex.predicted)
    assert extra == ({'mfma_fp64'} if name == 'ader_dense_f64' else set())
    assert ex.timed[ex.label] == best and extra <= set(pr.timed)
    assert pick.label in ex.predicted and not pick.timed
    assert 1 <= len(set(pr.timed) - extra) <= 3 and pr.label == min(pr.timed, key=pr.timed.get)
    assert pick.generator.get_launcher() and pr.generator.get_launcher()
    s_pick = ex.timed[pick.label] / best - 1
    s_pr = ex.timed[pr.label] / best - 1
    print(f'{name}: batch {batch}, {len(ex.timed)} candidates, best {ex.label} '
          f'{best * 1e3:.3f} ms; pick {pick.label} +{s_pick:.1%}; pruned {pr.label} '
          f'+{s_pr:.1%} ({len(pr.timed)} timed)')
    if name == 'ader_f64':
      assert s_pick <= 0.05 and s_pr <= 0.05
