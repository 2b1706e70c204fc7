"""Kernel names and launcher prototypes against the reference generator's own output.

``tests/golden/kernel_names.json`` holds, for every case of ``tests/golden/name_cases.py``, the
name (``kernel_`` + md5 prefix of the description string, reference
``tensorforge/generators/generator.py:349-362``), the launcher header (``generator.py:426-436``),
the call-site strings (``generator.py:445-492``) and the launch block that the reference
generator produced for it
(``tests/golden/make_kernel_names.py``). The same cases are built here with this package; a
caller that links launchers by name and declares them with the reference's header finds the same
symbols with the same parameters. Pointwise (``ElementwiseDescr``) names are not among them: the
reference hashes its statements' default object representation (its optree nodes define no
``__str__``), so its names for them change from run to run.
"""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, 'golden'))

import name_cases  # noqa: E402

with open(os.path.join(HERE, 'golden', 'kernel_names.json'), encoding='utf-8') as _f:
  GOLDEN = {r['label']: r for r in json.load(_f)}


def _generate(fp_name, descrs):
  import tensorforge
  assert os.path.abspath(tensorforge.__file__).startswith(REPO)
  from tensorforge.common.basic_types import FloatingPointType
  from tensorforge.common.context import Context
  from tensorforge.generators.generator import Generator
  fp = {'double': FloatingPointType.DOUBLE, 'float': FloatingPointType.FLOAT}[fp_name]
  gen = Generator(descrs, Context(arch='gfx950', backend='hip', fp_type=fp))
  gen.generate()
  return gen


def test_every_case_has_a_golden_record():
  labels = [label for label, _, _ in name_cases.cases()]
  assert sorted(labels) == sorted(GOLDEN)
  assert len(labels) >= 12


@pytest.mark.parametrize('case', name_cases.cases(), ids=lambda c: c[0])
def test_name_and_prototype_match_the_reference(case):
  label, fp_name, descrs = case
  gold = GOLDEN[label]
  assert gold['fp_type'] == fp_name
  gen = _generate(fp_name, descrs)
  assert gen.get_base_name() == gold['name']
  if gold['header'] is None:
    # the reference's own code generation stops on this case after naming it
    assert 'header_error' in gold
    return
  assert ' '.join(gen.get_header().split()) == ' '.join(gold['header'].split())


@pytest.mark.parametrize('case', name_cases.cases(), ids=lambda c: c[0])
def test_plain_launch_shape_matches_the_reference(case):
  """The plain emitter restates the reference's thread mapping: the same block (lanes per
  element, elements per block) and the same dynamic LDS size as the reference's launcher for
  gfx950 (each element's region padded to the vector unit's width plus the bank term)."""
  from tensorforge.common.basic_types import FloatingPointType
  from tensorforge.common.context import Context
  from tensorforge.generators.generator import Generator
  label, fp_name, descrs = case
  gold = GOLDEN[label]
  if gold.get('block') is None:
    pytest.skip('the reference generates no launcher for this case')
  fp = {'double': FloatingPointType.DOUBLE, 'float': FloatingPointType.FLOAT}[fp_name]
  gen = Generator(descrs, Context(arch='gfx950', backend='hip', fp_type=fp), emitter='plain')
  gen.generate()
  em = gen.get_emitted()
  assert list(em.block) == gold['block']
  words = em.schedule['lds_words_per_element']
  assert words * em.elements_per_block == gold['shared_words']


@pytest.mark.parametrize('case', name_cases.cases(), ids=lambda c: c[0])
def test_call_sites_match_the_reference(case):
  label, fp_name, descrs = case
  gold = GOLDEN[label]
  gen = _generate(fp_name, descrs)
  norm = lambda text: ' '.join(text.split())
  assert norm(gen.default_generate_call_site()) == norm(gold['default_call_site'])
  mats, offs = name_cases.call_site_maps(descrs)
  assert norm(gen.generate_call_site(mats, offs, 'numElements', 'flags', 'stream')) == \
      norm(gold['call_site'])


def test_catalog_header_exports_the_reference_names():
  """The prebuilt ADER and GEMM launchers are declared under the reference's names."""
  with open(os.path.join(REPO, 'include', 'tensorforge_catalog.h'), encoding='utf-8') as f:
    text = f.read()
  for label in ('ader_volume_o6_sparse', 'gemm_56x9_add'):
    assert f'launcher_{GOLDEN[label]["name"]}(' in text, label
