"""A launcher whose one-time setup is refused by the HIP runtime records the error and returns
without launching; the next, valid launcher in the same process must succeed and match the
oracle (HIP's last error is cleared where the refusal is recorded). The refused call is a host
API error (an LDS request over the workgroup limit); no kernel runs with it."""
import pytest

import launch_error_case as lec
from tensorforge_amd import catalog

from test_gpu_plain import _check

pytestmark = pytest.mark.gpu


def test_refused_setup_does_not_leak_into_next_launcher(device):
  import torch
  from tensorforge_amd.runtime.loader import load_kernels, runtime
  from tensorforge_amd.tune.measure import synthetic_buffers
  from tensorforge_amd.device import launch_args
  g = lec.generator()
  probe = load_kernels(lec.so_path(), [g])[0]
  n = 64
  bufs = synthetic_buffers(g, n, device)
  bufs.pop('__keep__')
  print('probe launcher', probe.name, 'LDS request', lec.OVER_LIMIT_LDS, flush=True)
  with pytest.raises(RuntimeError, match='launcher_kernel_launch_error_probe'):
    probe(*launch_args(g, bufs, n))
  torch.cuda.synchronize()
  assert runtime().tensorforge_last_error() == 0
  for name in ('gemm_f64_tuned', 'ader_f64_tuned', 'gemm_f64_plain'):
    entry, kernel = catalog.load(name)
    print('valid launcher after the refused one:', name, flush=True)
    _check(entry, kernel, device, 100, seed=5)
  assert runtime().tensorforge_last_error() == 0
