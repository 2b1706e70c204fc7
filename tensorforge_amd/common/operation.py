"""Pointwise operations of the optree description language.

Same member names as the reference's ``Operation`` (``tensorforge/common/operation.py:6-56``).
There, the hyperbolic functions share the values of the circular ones (``SINH = 100`` next to
``SIN = 100`` ...), which makes them aliases under Python's ``Enum``: ``sinh(x)`` is generated as
``sin(x)``. Here every member has a value of its own, so ``sinh`` is ``sinh``.
"""
import enum


class Operation(enum.Enum):
  COPY = 0
  MIN = 1
  MAX = 2
  DIV = 3
  MUL = 4
  ADD = 5
  SUB = 6
  SQRT = 7
  POW = 8
  CBRT = 9
  MOD = 10
  NEG = 11
  RCP = 12
  RSQRT = 13
  RCBRT = 14
  ABS = 15
  CEIL = 30
  FLOOR = 31
  ROUND = 32
  SIGN = 33
  TRUNC = 34
  LOGP1 = 94
  EXPM1 = 95
  GAMMA = 96
  ERF = 97
  LOG = 98
  EXP = 99
  SIN = 100
  COS = 101
  TAN = 102
  ASIN = 103
  ACOS = 104
  ATAN = 105
  SINH = 110
  COSH = 111
  TANH = 112
  ASINH = 113
  ACOSH = 114
  ATANH = 115
  NOT = 1000
  AND = 1001
  OR = 1002
  XOR = 1003
  SHR = 1100
  SHL = 1101
  SHRS = 1102
  EQ = 2000
  NEQ = 2001
  LT = 2002
  LE = 2003
  GT = 2004
  GE = 2005

  def __str__(self):
    return self.name.lower()


UNARY_MATH = {
  Operation.SQRT, Operation.CBRT, Operation.RSQRT, Operation.RCBRT, Operation.ABS,
  Operation.CEIL, Operation.FLOOR, Operation.ROUND, Operation.SIGN, Operation.TRUNC,
  Operation.LOGP1, Operation.EXPM1, Operation.GAMMA, Operation.ERF, Operation.LOG, Operation.EXP,
  Operation.SIN, Operation.COS, Operation.TAN, Operation.ASIN, Operation.ACOS, Operation.ATAN,
  Operation.SINH, Operation.COSH, Operation.TANH, Operation.ASINH, Operation.ACOSH,
  Operation.ATANH,
}
BINARY_MATH = {Operation.MIN, Operation.MAX, Operation.POW}
ARITHMETIC = {Operation.ADD, Operation.SUB, Operation.MUL, Operation.DIV, Operation.MOD}
COMPARISON = {Operation.EQ, Operation.NEQ, Operation.LT, Operation.LE, Operation.GT, Operation.GE}
BITWISE = {Operation.AND, Operation.OR, Operation.XOR, Operation.SHR, Operation.SHL,
           Operation.SHRS}
