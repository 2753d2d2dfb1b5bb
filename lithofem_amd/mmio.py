"""MatrixMarket exchange of the assembled system, in the reference's format.

The reference's solve driver exports the eliminated complex system as
``<prefix>.mtx`` (``%%MatrixMarket matrix coordinate complex general``,
1-based row-major entries of the sorted-column CSR, 17 significant digits)
and ``<prefix>.rhs.mtx`` (``matrix array complex general``, n x 1), and
imports a solution vector in the array format
(solver/lithofem_solve.cpp:254-307).  These helpers write and read the same
files, so a system assembled here can be compared with one assembled by an
MFEM-equipped build of the reference, or solved elsewhere.

Text I/O of this kind runs at roughly a microsecond per entry; it is meant
for parity checks and exchange, not for the hot path.
"""

from __future__ import annotations

import numpy as np

_COORD = "%%MatrixMarket matrix coordinate complex general"
_ARRAY = "%%MatrixMarket matrix array complex general"


def _chunks(n, size=1 << 20):
    for a in range(0, n, size):
        yield a, min(n, a + size)


def export_system(prefix: str, I, J, vre, vim, b=None) -> None:
    """Write ``prefix.mtx`` (and ``prefix.rhs.mtx`` when b is given) as the
    reference's export_system does."""
    I = np.asarray(I, dtype=np.int64)
    J = np.asarray(J, dtype=np.int64)
    vre = np.asarray(vre, dtype=np.float64)
    vim = np.asarray(vim, dtype=np.float64)
    n = len(I) - 1
    nnz = len(J)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(I))
    with open(prefix + ".mtx", "w") as f:
        f.write(f"{_COORD}\n{n} {n} {nnz}\n")
        for a, z in _chunks(nnz):
            f.write("".join("%d %d %.17g %.17g\n" % t for t in
                            zip((rows[a:z] + 1).tolist(), (J[a:z] + 1).tolist(),
                                vre[a:z].tolist(), vim[a:z].tolist())))
    if b is not None:
        write_vector(prefix + ".rhs.mtx", b)


def write_vector(path: str, x) -> None:
    """n x 1 complex array file (the reference's rhs export / solution
    import format)."""
    x = np.asarray(x, dtype=np.complex128)
    with open(path, "w") as f:
        f.write(f"{_ARRAY}\n{len(x)} 1\n")
        for a, z in _chunks(len(x)):
            f.write("".join("%.17g %.17g\n" % t for t in
                            zip(x[a:z].real.tolist(), x[a:z].imag.tolist())))


def _body(path: str):
    with open(path) as f:
        header = f.readline().strip()
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        sizes = [int(t) for t in line.split()]
        data = np.loadtxt(f, dtype=np.float64, ndmin=2)
    return header, sizes, data


def read_system(path: str):
    """Read a coordinate complex file into the sorted-column CSR
    (I, J, vre, vim); duplicate entries are summed."""
    header, sizes, data = _body(path)
    if header.lower() != _COORD.lower():
        raise ValueError(f"{path}: expected '{_COORD}', got '{header}'")
    n, m, nnz = sizes
    if n != m or data.shape[0] != nnz:
        raise ValueError(f"{path}: size line {sizes} does not match {data.shape[0]} entries")
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    key = r * n + c
    order = np.argsort(key, kind="stable")
    key, vre, vim = key[order], data[order, 2], data[order, 3]
    uk, start = np.unique(key, return_index=True)
    vre = np.add.reduceat(vre, start)
    vim = np.add.reduceat(vim, start)
    rows = uk // n
    I = np.zeros(n + 1, dtype=np.int64)
    np.add.at(I, rows + 1, 1)
    return np.cumsum(I).astype(np.int32), (uk % n).astype(np.int32), vre, vim


def read_vector(path: str) -> np.ndarray:
    """n x 1 complex array file -> complex vector (the reference's
    import_vector)."""
    header, sizes, data = _body(path)
    if header.lower() != _ARRAY.lower():
        raise ValueError(f"{path}: expected '{_ARRAY}', got '{header}'")
    if sizes[1] != 1 or data.shape[0] != sizes[0]:
        raise ValueError(f"{path}: expected {sizes[0]} x 1 entries, got {data.shape}")
    return data[:, 0] + 1j * data[:, 1]
