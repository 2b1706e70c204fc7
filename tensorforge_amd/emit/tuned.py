"""Placeholder replaced below."""


def unsupported_reason(program, context, schedule):
  return 'tuned emitter not available'


def emit(program, context, base_name, schedule):
  raise NotImplementedError
