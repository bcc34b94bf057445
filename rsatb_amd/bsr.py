"""Block-sparse row (BSR) operands in host memory.

Layout (the contract rsatb's matrices follow, rsatb/pybind_rsatb.cpp:636-668 and
rsatb/calc/test/test_preconditioner.cpp:37-80): uniform b x b blocks, CSR over
block rows with int64 row pointers and int32 block-column indices, columns
ascending within a row, entries row-major inside each block.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class BSR:
    n_block_rows: int
    block_dim: int
    row_ptr: np.ndarray  # int64 [nbr + 1]
    col_ind: np.ndarray  # int32 [nnzb]
    values: np.ndarray   # [nnzb, b, b] float64 or complex128 (float32 / complex64 widened on upload)

    @property
    def nnzb(self) -> int:
        return int(self.col_ind.shape[0])

    @property
    def n_rows(self) -> int:
        return self.n_block_rows * self.block_dim

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.values)

    def nbytes(self) -> int:
        return self.values.nbytes + self.col_ind.nbytes + self.row_ptr.nbytes

    @staticmethod
    def from_blocks(nbr: int, b: int, rows: np.ndarray, cols: np.ndarray, blocks: np.ndarray) -> "BSR":
        """Assembles blocks given as (row, col, b x b block) triples.

        Duplicate (row, col) pairs are summed -- the Gamma-point image sum
        H(k=0) = sum_R H(R). The result is canonical: rows in order, columns
        ascending.
        """
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        blocks = np.asarray(blocks)
        if blocks.ndim != 3 or blocks.shape[1:] != (b, b):
            raise ValueError("blocks must have shape [n, b, b]")
        if rows.size and (rows.min() < 0 or rows.max() >= nbr or cols.min() < 0 or cols.max() >= nbr):
            raise ValueError("block index out of range")
        key = rows * nbr + cols
        order = np.argsort(key, kind="stable")
        key_s = key[order]
        if key_s.size:
            first = np.ones(key_s.size, dtype=bool)
            first[1:] = key_s[1:] != key_s[:-1]
            starts = np.flatnonzero(first)
            vals = np.add.reduceat(blocks[order], starts, axis=0) if starts.size else blocks[:0]
            ukey = key_s[starts]
        else:
            vals = blocks[:0]
            ukey = key_s
        urow = ukey // nbr
        ucol = ukey % nbr
        row_ptr = np.zeros(nbr + 1, dtype=np.int64)
        np.add.at(row_ptr, urow + 1, 1)
        row_ptr = np.cumsum(row_ptr).astype(np.int64)
        return BSR(nbr, b, row_ptr, ucol.astype(np.int32), np.ascontiguousarray(vals))

    def block_rows(self) -> np.ndarray:
        """Block-row index of every stored block."""
        return np.repeat(np.arange(self.n_block_rows, dtype=np.int64), np.diff(self.row_ptr))

    def conj_transpose(self) -> "BSR":
        """A^H as its own BSR (what the sandwich's Z^H factor is stored as)."""
        r = self.block_rows()
        blocks = np.conj(np.transpose(self.values, (0, 2, 1)))
        return BSR.from_blocks(self.n_block_rows, self.block_dim, self.col_ind.astype(np.int64), r, blocks)

    def astype(self, dtype) -> "BSR":
        return BSR(self.n_block_rows, self.block_dim, self.row_ptr, self.col_ind, self.values.astype(dtype))

    def to_dense(self) -> np.ndarray:
        n, b = self.n_rows, self.block_dim
        out = np.zeros((n, n), dtype=self.values.dtype)
        r = self.block_rows()
        for s in range(self.nnzb):
            i, j = int(r[s]), int(self.col_ind[s])
            out[i * b:(i + 1) * b, j * b:(j + 1) * b] += self.values[s]
        return out

    def to_coo(self):
        """Global COO triplets (dump_operator_coo, rsatb/pybind_rsatb.cpp:612-670)."""
        b = self.block_dim
        r = self.block_rows()
        ii, jj = np.meshgrid(np.arange(b), np.arange(b), indexing="ij")
        rows = (r[:, None, None] * b + ii[None]).reshape(-1)
        cols = (self.col_ind.astype(np.int64)[:, None, None] * b + jj[None]).reshape(-1)
        return rows, cols, self.values.reshape(-1)

    def check(self) -> None:
        if self.row_ptr.shape != (self.n_block_rows + 1,) or self.row_ptr[0] != 0:
            raise ValueError("bad row_ptr")
        if np.any(np.diff(self.row_ptr) < 0) or self.row_ptr[-1] != self.nnzb:
            raise ValueError("row_ptr must be non-decreasing and end at nnzb")
        if self.values.shape != (self.nnzb, self.block_dim, self.block_dim):
            raise ValueError("values must be [nnzb, b, b]")
