"""Write ``kernel_names.json``: the reference generator's kernel name and launcher header for
every case of ``name_cases.py``.

Run in a checkout that holds the reference next to this repository (it imports the reference's
``tensorforge`` package from ``/root/reference``; nothing of the reference is copied, only the
names and prototypes it produces):

    python tests/golden/make_kernel_names.py

The kernel name is ``kernel_`` + the first 10 hex digits of the md5 of the description string
(reference ``tensorforge/generators/generator.py:349-362``); the header is the launcher prototype
(``generator.py:426-436``); the call sites are ``default_generate_call_site()`` and
``generate_call_site`` with ``<alias>_ptr`` / ``<alias>_off`` names (``generator.py:445-492``);
the launch shape is the block (lanes per element, elements per block)
and the dynamic LDS size in words that the reference's launcher passes
(``generator.py:438-492``, chosen by ``RegmaxBlockPolicy`` and the LDS allocator), for the
reference's own ``gfx950`` entry (``common/vm/hw_descr_db.yml``).
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = '/root/reference'


def main():
  sys.path.insert(0, REFERENCE)
  import tensorforge
  assert os.path.abspath(tensorforge.__file__).startswith(REFERENCE), tensorforge.__file__
  from tensorforge.common.basic_types import FloatingPointType
  from tensorforge.common.context import Context
  from tensorforge.generators.generator import Generator
  import name_cases

  fp = {'double': FloatingPointType.DOUBLE, 'float': FloatingPointType.FLOAT}
  records = []
  for label, fp_name, descrs in name_cases.cases():
    ctx = Context(arch='gfx950', backend='hip', fp_type=fp[fp_name])
    gen = Generator(descrs, ctx)
    gen.register()
    rec = {'label': label, 'fp_type': fp_name, 'name': gen.get_base_name(), 'header': None,
           'default_call_site': gen.default_generate_call_site()}
    mats, offs = name_cases.call_site_maps(descrs)
    try:
      rec['call_site'] = gen.generate_call_site(mats, offs, 'numElements', 'flags', 'stream')
    except Exception as exc:
      rec['call_site_error'] = f'{type(exc).__name__}: {exc}'
    try:
      gen.generate()
      rec['header'] = gen.get_header()
      launcher = gen.get_launcher()
      block = re.search(r'dim3 block\s*\(([^)]*)\)', launcher)
      words = re.search(r'(\d+) \* sizeof\(', launcher)
      rec['block'] = [int(v) for v in block.group(1).split(',')] if block else None
      rec['shared_words'] = int(words.group(1)) if words else 0
    except Exception as exc:  # the name is still pinned; say why the prototype is not
      rec['header_error'] = f'{type(exc).__name__}: {exc}'
    records.append(rec)
    print(label, rec['name'], 'header' if rec['header'] else rec.get('header_error'))
  with open(os.path.join(HERE, 'kernel_names.json'), 'w', encoding='utf-8') as f:
    json.dump(records, f, indent=1)
    f.write('\n')


if __name__ == '__main__':
  main()
