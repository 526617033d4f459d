"""File formats of the reference's data path.

* fvecs: repeated ``[int32 d][d x float32]`` (src/utils.h:204-237).
* results ``.bin``: ``int64 I[nq*k]`` then ``float32 D[nq*k]``, no header
  (src/utils.h:585-593, 610-618).
"""
from __future__ import annotations

import numpy as np


def read_fvecs(path: str, leading_word: bool = False) -> np.ndarray:
    """Read an fvecs file.  ``leading_word`` skips one extra int32 at the
    start (the reference's ``benchmarks/test`` files carry one)."""
    raw = np.fromfile(path, dtype=np.int32)
    if leading_word:
        raw = raw[1:]
    if raw.size == 0:
        return np.zeros((0, 0), np.float32)
    d = int(raw[0])
    rec = d + 1
    if raw.size % rec:
        raise ValueError(f"{path}: size {raw.size * 4} B is not a multiple of {rec * 4} B")
    rows = raw.reshape(-1, rec)
    if not np.all(rows[:, 0] == d):
        raise ValueError(f"{path}: inconsistent record dimensions")
    return rows[:, 1:].copy().view(np.float32)


def write_fvecs(path: str, x: np.ndarray) -> None:
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    out = np.empty((n, d + 1), np.int32)
    out[:, 0] = d
    out[:, 1:] = x.view(np.int32)
    out.tofile(path)


def write_results(path: str, I: np.ndarray, D: np.ndarray) -> None:
    with open(path, "wb") as f:
        np.ascontiguousarray(I, np.int64).tofile(f)
        np.ascontiguousarray(D, np.float32).tofile(f)


def read_results(path: str, nq: int, k: int):
    """Inverse of write_results (src/utils.h:610-618); the file must hold
    exactly nq * k results."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size != nq * k * 12:
        raise ValueError(f"{path}: {raw.size} B, expected {nq * k * 12} B for {nq} x {k} results")
    I = raw[: nq * k * 8].view(np.int64).reshape(nq, k).copy()
    D = raw[nq * k * 8: nq * k * 12].view(np.float32).reshape(nq, k).copy()
    return I, D
