"""``import tensorforge`` for code written against the reference package.

Every ``tensorforge.<module>`` import resolves to the same module object as
``tensorforge_amd.<module>``, so classes and enums are shared between the two names.
"""
import importlib
import importlib.abc
import importlib.util
import sys

import tensorforge_amd as _impl
from tensorforge_amd import *  # noqa: F401,F403


class _AliasFinder(importlib.abc.MetaPathFinder, importlib.abc.Loader):
  PREFIX = 'tensorforge.'

  def find_spec(self, fullname, path, target=None):
    if not fullname.startswith(self.PREFIX):
      return None
    return importlib.util.spec_from_loader(fullname, self)

  def create_module(self, spec):
    target = 'tensorforge_amd.' + spec.name[len(self.PREFIX):]
    module = importlib.import_module(target)
    self._original_spec = module.__spec__
    sys.modules[spec.name] = module
    return module

  def exec_module(self, module):
    # the import machinery stamps the alias spec on the shared module; put the real one back
    module.__spec__ = self._original_spec


if not any(isinstance(f, _AliasFinder) for f in sys.meta_path):
  sys.meta_path.insert(0, _AliasFinder())
__path__ = []
