"""Readers for the PETSc binary files held by the reference's tests (data only).

Mat: big-endian int64 header {1211216, M, N, nz}, M int64 row lengths, nz int64
column indices, nz big-endian complex128 values (PetscInt=int64, complex build).
Vec: {1211214, N} + N complex128.
"""
import numpy as np

MAT_ID = 1211216
VEC_ID = 1211214


def read_mat(path):
    with open(path, "rb") as f:
        buf = f.read()
    hdr = np.frombuffer(buf, ">i8", 4, 0)
    assert int(hdr[0]) == MAT_ID, hdr
    M, N, nz = int(hdr[1]), int(hdr[2]), int(hdr[3])
    off = 32
    rl = np.frombuffer(buf, ">i8", M, off).astype(np.int64)
    off += 8 * M
    col = np.frombuffer(buf, ">i8", nz, off).astype(np.int64)
    off += 8 * nz
    val = np.frombuffer(buf, ">c16", nz, off).astype(np.complex128)
    off += 16 * nz
    assert off == len(buf)
    rowptr = np.zeros(M + 1, np.int64)
    np.cumsum(rl, out=rowptr[1:])
    return M, N, rowptr, col, val


def read_vec(path):
    with open(path, "rb") as f:
        buf = f.read()
    hdr = np.frombuffer(buf, ">i8", 2, 0)
    assert int(hdr[0]) == VEC_ID, hdr
    n = int(hdr[1])
    v = np.frombuffer(buf, ">c16", n, 16).astype(np.complex128)
    assert 16 + 16 * n == len(buf)
    return v


def to_dense(M, N, rowptr, col, val):
    A = np.zeros((M, N), np.complex128)
    for r in range(M):
        for k in range(rowptr[r], rowptr[r + 1]):
            A[r, col[k]] += val[k]
    return A
