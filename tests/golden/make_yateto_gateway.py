"""Write ``yateto_gateway.json``: what the reference's yateto frontend produces through its
legacy gateway (``YatetoFrontend.add_linear_operation`` -> ``add_operation``, reference
``tensorforge/frontend/yateto.py:27-38,96-125,150-168,390-392``) for every case of
``yateto_cases.py``: the description strings (bounding boxes after alignment and views), the
call-site lines written into yateto's C++, the routine names and the prototypes the routine
writers return.

Run in a checkout that holds the reference next to this repository (it imports the reference's
``tensorforge`` from ``/root/reference``; only these strings are kept):

    python tests/golden/make_yateto_gateway.py
"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = '/root/reference'


def main():
  sys.path.insert(0, HERE)
  sys.path.insert(0, REFERENCE)
  import tensorforge
  assert os.path.abspath(tensorforge.__file__).startswith(REFERENCE), tensorforge.__file__
  import yateto_cases as yc
  out = {}
  for case in yc.CASES:
    rec = {}
    try:
      _, lines, cache, descrs = yc.run(tensorforge, case)
      rec['descrs'] = descrs
      rec['cpp'] = lines
      rec['routines'] = sorted(cache.routines)
      headers = {}
      with tempfile.TemporaryDirectory() as tmp:
        for name, writer in cache.routines.items():
          headers[name] = writer(name, os.path.join(tmp, 'routines.cpp'))
      rec['headers'] = headers
    except Exception as exc:
      rec['error'] = f'{type(exc).__name__}: {exc}'
    out[case] = rec
    print(case, rec.get('routines', rec.get('error')))
  with open(os.path.join(HERE, 'yateto_gateway.json'), 'w', encoding='utf-8') as f:
    json.dump(out, f, indent=1)
    f.write('\n')


if __name__ == '__main__':
  main()
