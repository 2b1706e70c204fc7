"""The CMake-built C++ consumer (tests/cmake_consumer): a contraction with fused pointwise
statements from a generated routine file, called through its C ABI from C++, checked inside the
program against a CPU loop (normwise per element <= 1e-12)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'cmake_consumer', '_build',
                   'consumer')


def test_cpp_consumer_runs():
  if not os.path.exists(EXE):
    pytest.skip('the consumer is built by __graft_entry__.build() (cmake was not available)')
  print(f'running {EXE}', flush=True)
  r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
  print(r.stdout, r.stderr, flush=True)
  assert r.returncode == 0, f'exit {r.returncode}: {r.stdout} {r.stderr}'
  assert 'max normwise error' in r.stdout
