"""Copy the C headers next to a consumer's sources (reference ``tensorforge/support.py:5-21``)."""
import os
import shutil


def copy_includes_and_aux(destination, name_as=None):
  if not os.path.isdir(destination):
    raise NotADirectoryError(f'there is not such dir as {destination}')
  src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include')
  trg = os.path.join(destination, name_as if name_as else 'include')
  if os.path.exists(trg):
    raise FileExistsError(f'a directory with a name ({name_as}) is exists in {destination}')
  shutil.copytree(src, trg)
