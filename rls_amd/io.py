"""Matrix input and grid-result output.

* ``loaddata`` follows loaddata (src/utils/loadmtx.jl:1-28): ``<data_dir>/<mname>.mtx`` read
  as Matrix Market, converted to complex128, with 32-bit sparse indices when the dimensions
  and entries fit.  A coordinate file gives a CSC matrix (symmetric, skew-symmetric and
  Hermitian storage expanded to the full matrix, as the Matrix Market reader does), an array
  file a dense array.  The reference looks in the package's ``data`` folder; here the folder
  is the argument, else $RLS_DATA_DIR, else ``data`` beside the package.
* ``save_grid_result`` writes a solve's output in natural grid order (the layout of
  reorder_to_natural_grid, src/core/recuGrid.jl:229-232, which the reference's drivers
  pass to their contour plots) as a NumPy ``.npz`` archive in the place of the drivers'
  Julia serialization files.
"""
from __future__ import annotations

import os
from pathlib import Path
from typing import Optional

import numpy as np


def _data_dir(data_dir: Optional[str]) -> Path:
    if data_dir is not None:
        return Path(data_dir)
    env = os.environ.get("RLS_DATA_DIR")
    if env:
        return Path(env)
    return Path(__file__).resolve().parent.parent / "data"


def loaddata(mname: str, data_dir: Optional[str] = None, use_int32_index: bool = True):
    """The matrix ``<data_dir>/<mname>.mtx`` as complex128 (loadmtx.jl:1-28)."""
    import scipy.io
    import scipy.sparse as sp
    path = _data_dir(data_dir) / (mname + ".mtx")
    if not path.is_file():
        raise FileNotFoundError(f"Matrix file not found at {path}")
    A = scipy.io.mmread(str(path))
    if sp.issparse(A):
        A = sp.csc_matrix(A, dtype=np.complex128)
        A.sum_duplicates()
        i32 = np.iinfo(np.int32).max
        nrow, ncol = A.shape
        max_row = int(A.indices.max()) if A.nnz else 0
        if use_int32_index and nrow <= i32 and ncol <= i32 and max_row <= i32 and len(A.indptr) <= i32:
            A.indices = A.indices.astype(np.int32)
            A.indptr = A.indptr.astype(np.int32)
        else:
            A.indices = A.indices.astype(np.int64)
            A.indptr = A.indptr.astype(np.int64)
        return A
    return np.asarray(A, dtype=np.complex128)


def save_grid_result(path, rec, sigma, **extra) -> None:
    """Write sigma (traversal order) and its natural-order grid to ``path`` (.npz): arrays
    ``grid`` (n_y x n_x, grid[iy, ix] at x[ix] + i y[iy]), ``x``, ``y``, ``Z``, ``sigma`` and
    the keyword arrays given."""
    from . import reorder_to_natural_grid
    sigma = np.asarray(sigma)
    x = rec.x_st + rec.dx * np.arange(rec.n_x)
    y = rec.y_st + rec.dy * np.arange(rec.n_y)
    np.savez(path, grid=reorder_to_natural_grid(rec, sigma), x=x, y=y, Z=np.asarray(rec.Z), sigma=sigma,
             **{k: np.asarray(v) for k, v in extra.items()})
