"""yateto's legacy gateway (``YatetoFrontend.add_linear_operation`` -> ``add_operation``,
reference ``tensorforge/frontend/yateto.py:27-38,96-125,150-168,390-392``) on yateto-shaped
objects (``tests/golden/yateto_cases.py``): per case, the description strings (bounding boxes
after leading-dimension alignment with ``arch.alignedLower/alignedUpper``, ``alignedStride()``
and ``MemoryLayoutView`` offsets), the launcher call written into yateto's C++, the routine names
and the prototypes the routine writers return must equal what the reference's frontend produced
(``tests/golden/yateto_gateway.json``, ``tests/golden/make_yateto_gateway.py``). This package's
writer declares the prototype ``extern "C"``; the rest of the text is identical."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, 'golden'))

import yateto_cases as yc  # noqa: E402

with open(os.path.join(HERE, 'golden', 'yateto_gateway.json'), encoding='utf-8') as _f:
  GOLDEN = json.load(_f)


def _package():
  import tensorforge
  assert os.path.abspath(tensorforge.__file__).startswith(REPO)
  return tensorforge


@pytest.mark.parametrize('case', list(yc.CASES))
def test_legacy_gateway_matches_the_reference(case, tmp_path):
  want = GOLDEN[case]
  assert 'error' not in want, want.get('error')
  fe, lines, cache, descrs = yc.run(_package(), case)
  assert descrs == want['descrs']
  assert lines == want['cpp']
  assert sorted(cache.routines) == want['routines']
  for name, writer in cache.routines.items():
    header = writer(name, str(tmp_path / 'routines.cpp'))
    assert header.startswith('extern "C" ')
    assert header[len('extern "C" '):] == want['headers'][name]
  # the routine file holds the kernel and its launcher, and compiles on its own for gfx950
  text = (tmp_path / 'routines.cpp').read_text()
  for name in cache.routines:
    assert f'void launcher_{name}(' in text


def test_alignment_widens_only_aligned_strides():
  _, _, _, aligned = yc.run(_package(), 'aligned_window')
  _, _, _, plain = yc.run(_package(), 'unaligned_window')
  assert '({0..56}×{0..9})' in aligned[0] and '({3..51}×{0..9})' in plain[0]


def test_view_operand_reads_its_window_of_the_storage():
  _, _, _, d = yc.run(_package(), 'view_operand')
  assert '56×18(56×18)' in d[0] and '({0..56}×{9..18})' in d[0]
