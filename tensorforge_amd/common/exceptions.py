"""Error types raised by description checking and kernel generation.

Mirrors the reference's ``tensorforge/common/exceptions.py:1-6`` so callers that catch
``GenerationError`` / ``InternalError`` keep working.
"""


class GenerationError(Exception):
  """A description cannot be turned into a kernel (bad shapes, bounding boxes, patterns)."""


class InternalError(Exception):
  """An inconsistency inside the emitter; indicates a bug rather than a user error."""
