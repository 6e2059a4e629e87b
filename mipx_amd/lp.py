"""LP container in the reference's ranged form, held as numpy arrays.

Mirrors the fields of `mipx::LpProblem` that the PDLP path reads
(include/mipx/lp_problem.h:14-64): obj, col_lower/upper, row_lower/upper,
a CSR matrix with sorted column indices per row, sense and obj_offset.
"""
from dataclasses import dataclass, field

import ctypes as C
import numpy as np

from ._abi import MpxLp, c_double_p, c_int32_p

INF = float("inf")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


@dataclass
class LpProblem:
    num_rows: int
    num_cols: int
    row_starts: np.ndarray
    col_indices: np.ndarray
    values: np.ndarray
    obj: np.ndarray
    col_lower: np.ndarray
    col_upper: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    sense: int = 0          # 0 minimize, 1 maximize
    obj_offset: float = 0.0
    name: str = ""
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        self.row_starts = _i32(self.row_starts)
        self.col_indices = _i32(self.col_indices)
        self.values = _f64(self.values)
        self.obj = _f64(self.obj)
        self.col_lower = _f64(self.col_lower)
        self.col_upper = _f64(self.col_upper)
        self.row_lower = _f64(self.row_lower)
        self.row_upper = _f64(self.row_upper)
        assert self.row_starts.shape == (self.num_rows + 1,)
        assert self.obj.shape == (self.num_cols,)

    @property
    def nnz(self):
        return int(self.row_starts[-1]) if self.num_rows > 0 else 0

    def as_c(self):
        """MpxLp view (the arrays stay owned by this object)."""
        lp = MpxLp()
        lp.num_rows = self.num_rows
        lp.num_cols = self.num_cols
        lp.nnz = self.nnz
        lp.sense = int(self.sense)
        lp.row_starts = self.row_starts.ctypes.data_as(c_int32_p)
        lp.col_indices = self.col_indices.ctypes.data_as(c_int32_p)
        lp.values = self.values.ctypes.data_as(c_double_p)
        lp.obj = self.obj.ctypes.data_as(c_double_p)
        lp.col_lower = self.col_lower.ctypes.data_as(c_double_p)
        lp.col_upper = self.col_upper.ctypes.data_as(c_double_p)
        lp.row_lower = self.row_lower.ctypes.data_as(c_double_p)
        lp.row_upper = self.row_upper.ctypes.data_as(c_double_p)
        lp.obj_offset = float(self.obj_offset)
        return lp

    def dense(self):
        a = np.zeros((self.num_rows, self.num_cols))
        for i in range(self.num_rows):
            s, e = self.row_starts[i], self.row_starts[i + 1]
            a[i, self.col_indices[s:e]] += self.values[s:e]
        return a

    def with_col_bounds(self, lo, hi):
        return LpProblem(self.num_rows, self.num_cols, self.row_starts, self.col_indices,
                         self.values, self.obj, lo, hi, self.row_lower, self.row_upper,
                         self.sense, self.obj_offset, self.name)


def from_triplets(num_rows, num_cols, triplets, obj, col_lower, col_upper,
                  row_lower, row_upper, sense=0, obj_offset=0.0, name=""):
    """Build CSR like SparseMatrix(rows, cols, triplets): sort by (row, col),
    sum duplicates and drop entries that sum to zero
    (src/lp/sparse_matrix.cpp:12-73)."""
    acc = {}
    for (i, j, v) in triplets:
        acc[(int(i), int(j))] = acc.get((int(i), int(j)), 0.0) + float(v)
    keys = sorted(k for k, v in acc.items() if v != 0.0)
    rs = np.zeros(num_rows + 1, dtype=np.int32)
    for (i, _j) in keys:
        rs[i + 1] += 1
    rs = np.cumsum(rs).astype(np.int32)
    ci = np.array([j for (_i, j) in keys], dtype=np.int32)
    va = np.array([acc[k] for k in keys], dtype=np.float64)
    return LpProblem(num_rows, num_cols, rs, ci, va, obj, col_lower, col_upper,
                     row_lower, row_upper, sense, obj_offset, name)


def from_dense(a, obj, col_lower, col_upper, row_lower, row_upper, sense=0, obj_offset=0.0):
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    trips = [(i, j, a[i, j]) for i in range(m) for j in range(n) if a[i, j] != 0.0]
    return from_triplets(m, n, trips, obj, col_lower, col_upper, row_lower, row_upper,
                         sense, obj_offset)
