"""Build support for tests/test_gpu_autotune.py: every candidate form of the chains the test
tunes, compiled into tests/autotune_build/_build (one shared object per chain).

    python tests/autotune_build.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
  if p not in sys.path:
    sys.path.insert(0, p)

BUILD_DIR = os.path.join(HERE, 'autotune_build', '_build')
# random chains of the cost model's workload drawn outside the committed database's range
# (seed 2026, indices 40000 and up), and the order-6 volume chain
SEED, INDICES = 2026, (40000, 40001, 40002)


def chains():
  from tensorforge_amd import workloads as W
  from tensorforge_amd.common.basic_types import FloatingPointType
  from tensorforge_amd.tune.sweep import chain_descriptions
  out = []
  for i in INDICES:
    descrs, fp, batch, info = chain_descriptions(SEED, i)
    out.append((f'random_{i}', descrs, fp, batch))
  w = W.ader_volume(FloatingPointType.DOUBLE)
  out.append(('ader_f64', w.descrs, w.fp_type, 262144))
  # dense K_d: the FP64 matrix-core form is admitted as well (timed, never predicted)
  w = W.ader_volume(FloatingPointType.DOUBLE, sparse_k=False)
  out.append(('ader_dense_f64', w.descrs, w.fp_type, 262144))
  return out


def build(directory=BUILD_DIR):
  from tensorforge_amd.tune.autotune import prepare
  return [prepare(d, fp, batch, directory, force=True) for _, d, fp, batch in chains()]


if __name__ == '__main__':
  print('\n'.join(build()))
