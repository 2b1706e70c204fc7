"""The runtime library's copy probe (the HBM bandwidth roof, restating the reference's
benchmarks/bandwidth/global.cpp:30-39, To[i] = From[i]): exact copies at sizes that end inside
a workgroup's share, inside a lane's chunk group and on a boundary; odd sizes are refused."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n', [2, 6, 2048, 2050, 4096 + 10, (1 << 20) + 514])
def test_copy_probe_is_exact(n, device):
  import torch
  from tensorforge_amd.runtime.loader import runtime
  rt = runtime()
  src = torch.randn(n + 64, dtype=torch.float64, device=device)
  dst = torch.full((n + 64,), 7.0, dtype=torch.float64, device=device)
  stream = torch.cuda.current_stream()
  print('copy probe n =', n, flush=True)
  assert rt.tensorforge_copy(dst.data_ptr(), src.data_ptr(), n, stream.cuda_stream) == 0
  torch.cuda.synchronize()
  assert torch.equal(dst[:n], src[:n])
  assert bool((dst[n:] == 7.0).all())


def test_copy_probe_refuses_odd_sizes(device):
  import torch
  from tensorforge_amd.runtime.loader import runtime
  rt = runtime()
  src = torch.zeros(8, dtype=torch.float64, device=device)
  assert rt.tensorforge_copy(src.data_ptr(), src.data_ptr(), 7, None) != 0
