"""yateto-shaped objects and the legacy-gateway cases whose results are pinned by
``yateto_gateway.json`` (written by ``make_yateto_gateway.py`` with the reference's frontend,
read back by tests/test_yateto_gateway.py with this package's).

The stand-ins carry exactly what the frontends read from yateto (``frontend/yateto.py:19-130,
170-330`` of the reference): a term's ``name``, ``indices``, ``memoryLayout``, ``eqspp``,
``is_temporary``, ``is_compute_constant``, ``addressing``, ``values``, ``datatype``; a layout's
``bbox()``, ``alignedStride()``, ``entries(rows, cols)`` (sparse layouts), ``storage()`` and
``relidx()`` (``MemoryLayoutView``), dispatched on the class name as the frontends do; the
architecture's ``name``, ``backend``, ``typename``, ``alignedLower`` / ``alignedUpper``.
No package is imported here: the frontend under test is passed in.
"""


class DenseMemoryLayout:
  def __init__(self, shape, aligned=False, bbox=None):
    self._shape = tuple(shape)
    self._bbox = [range(lo, hi) for lo, hi in (bbox or [(0, s) for s in shape])]
    self._aligned = aligned

  def bbox(self):
    return list(self._bbox)

  def alignedStride(self):
    return self._aligned

  def shape(self):
    return self._shape


class CSCMemoryLayout:
  """A sparse matrix layout: the stored (row, column) entries, column-major."""

  def __init__(self, shape, entries):
    self._shape = tuple(shape)
    self._entries = sorted(entries, key=lambda rc: (rc[1], rc[0]))

  def bbox(self):
    return [range(0, s) for s in self._shape]

  def alignedStride(self):
    return False

  def entries(self, rows, cols):
    return [(r, c) for r, c in self._entries if r in rows and c in cols]


class MemoryLayoutView:
  """A view at ``offset`` into a larger storage layout."""

  def __init__(self, storage, offset):
    self._storage = storage
    self._offset = tuple(offset)

  def storage(self):
    return self._storage

  def relidx(self, idx):
    return tuple(i + o for i, o in zip(idx, self._offset))

  def alignedStride(self):
    return self._storage.alignedStride()

  def bbox(self):
    return self._storage.bbox()


class EqSpp:
  def __init__(self, bounds):
    self._bounds = [tuple(b) for b in bounds]

  def nnzbounds(self):
    return list(self._bounds)


class Term:
  def __init__(self, name, indices, layout, bounds, temporary=False, constant=False,
               addressing=None, values=None, datatype=None):
    self.name = name
    self.indices = indices
    self.memoryLayout = layout
    self.eqspp = EqSpp(bounds)
    self.is_temporary = temporary
    self.is_compute_constant = constant
    self.addressing = addressing
    self.values = values
    self.datatype = datatype

  def __eq__(self, other):
    return isinstance(other, Term) and other.name == self.name

  def __hash__(self):
    return hash(self.name)


class Arch:
  """gfx950, FP64, 8-entry (64-byte) alignment of leading dimensions."""
  name = 'gfx950'
  backend = 'hip'
  typename = 'double'
  align = 8

  def alignedLower(self, x):
    return x - x % self.align

  def alignedUpper(self, x):
    return x + (-x) % self.align


class Cpp:
  """Collects the lines a frontend writes into yateto's generated C++."""

  def __init__(self):
    self.lines = []

  def __call__(self, line):
    self.lines.append(line)


class RoutineCache:
  def __init__(self):
    self.routines = {}

  def addRoutine(self, name, writer):
    self.routines[name] = writer


PAIR = [[0, -1], [-1, 1]]
IDENT = [[0, 1], [0, 1]]
STAR = sorted([(r, c) for r in range(3) for c in range(6, 9)] +
              [(3, 6), (3, 7), (4, 7), (4, 8), (5, 6), (5, 8)] +
              [(6, 0), (6, 3), (6, 5), (7, 3), (7, 1), (7, 4), (8, 5), (8, 4), (8, 2)],
              key=lambda rc: (rc[1], rc[0]))


def _k_entries(b, seed):
  """A fixed pseudo-random ~40 % pattern (linear congruential, no library)."""
  out, x = [], seed
  for c in range(b):
    for r in range(b):
      x = (1103515245 * x + 12345) % (1 << 31)
      if x % 5 < 2:
        out.append((r, c))
  return out


def gemm_ptr(fe):
  """C_ij += A_ik B_kj: A compute-constant (shared), B and C per element (pointer-based, as
  yateto's non-constant, non-temporary terms)."""
  A = Term('A', 'ik', DenseMemoryLayout((56, 56), aligned=True), [(0, 55), (0, 55)], constant=True)
  B = Term('B', 'kj', DenseMemoryLayout((56, 9), aligned=True), [(0, 55), (0, 8)])
  C = Term('C', 'ij', DenseMemoryLayout((56, 9), aligned=True), [(0, 55), (0, 8)])
  fe.add_linear_operation(C, [A, B], PAIR, IDENT, True)


def aligned_window(fe):
  """Leading-dimension alignment: the destination's nonzero rows 3..50 of a 64-row aligned
  storage widen to the aligned window [0, 56); the shared operand's rows with them."""
  A = Term('Aw', 'ik', DenseMemoryLayout((64, 20), aligned=True), [(3, 50), (0, 19)],
           constant=True)
  B = Term('Bw', 'kj', DenseMemoryLayout((20, 9), aligned=True), [(0, 19), (0, 8)])
  C = Term('Cw', 'ij', DenseMemoryLayout((64, 9), aligned=True), [(3, 50), (0, 8)])
  fe.add_linear_operation(C, [A, B], PAIR, IDENT, False)


def unaligned_window(fe):
  """The same window without an aligned stride: the nonzero bounds are kept as they are."""
  A = Term('Au', 'ik', DenseMemoryLayout((64, 20)), [(3, 50), (0, 19)], constant=True)
  B = Term('Bu', 'kj', DenseMemoryLayout((20, 9)), [(0, 19), (0, 8)])
  C = Term('Cu', 'ij', DenseMemoryLayout((64, 9)), [(3, 50), (0, 8)])
  fe.add_linear_operation(C, [A, B], PAIR, IDENT, True)


def ader_chain(fe):
  """The volume chain through the legacy gateway: a temporary per term (yateto's ``_tmp``
  terms), sparse star matrices and sparse shared K (CSC layouts)."""
  Q = Term('Q', 'kp', DenseMemoryLayout((56, 9), aligned=True), [(0, 55), (0, 8)],
           addressing=None)
  Qp = Term('Qp', 'lp', DenseMemoryLayout((56, 9), aligned=True), [(0, 55), (0, 8)])
  for d in range(3):
    S = Term(f'S{d}', 'qp', CSCMemoryLayout((9, 9), STAR), [(0, 8), (0, 8)])
    K = Term(f'K{d}', 'lk', CSCMemoryLayout((56, 56), _k_entries(56, 7 + d)), [(0, 55), (0, 55)],
             constant=True)
    T = Term(f'_tmp{d}', 'kp', DenseMemoryLayout((56, 9), aligned=True), [(0, 55), (0, 8)],
             temporary=True)
    Qv = Term('Q', 'kq', DenseMemoryLayout((56, 9), aligned=True), [(0, 55), (0, 8)])
    fe.add_linear_operation(T, [Qv, S], PAIR, IDENT, False)
    fe.add_linear_operation(Qp, [K, T], PAIR, IDENT, d > 0)
  del Q


def view_operand(fe):
  """An operand that is a view into a larger storage (columns 9..17 of a 56 x 18 array)."""
  A = Term('Av', 'ik', DenseMemoryLayout((56, 56), aligned=True), [(0, 55), (0, 55)],
           constant=True)
  big = DenseMemoryLayout((56, 18), aligned=True)
  B = Term('Bv', 'kj', MemoryLayoutView(big, (0, 9)), [(0, 55), (0, 8)])
  C = Term('Cv', 'ij', DenseMemoryLayout((56, 9), aligned=True), [(0, 55), (0, 8)])
  fe.add_linear_operation(C, [A, B], PAIR, IDENT, True)


CASES = {'gemm_ptr': gemm_ptr, 'aligned_window': aligned_window,
         'unaligned_window': unaligned_window, 'ader_chain': ader_chain,
         'view_operand': view_operand}


def run(package, case):
  """Drive ``package``'s yateto frontend through the legacy gateway for ``case``; returns the
  frontend, the C++ lines it wrote, the routine cache, and the description strings."""
  Frontend = package.get_routine_generator(None)
  fe = Frontend(Arch())
  CASES[case](fe)
  descrs = [str(d) for d in fe.generator._descr_list]
  cpp, cache = Cpp(), RoutineCache()
  fe.generate(cpp, cache)
  return fe, cpp.lines, cache, descrs
