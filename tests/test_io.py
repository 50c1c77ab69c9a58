"""Matrix Market input (loaddata, loadmtx.jl:1-28) and the natural-order result archive."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import rls_amd as R
from rls_amd import io as RIO


def test_loaddata_coordinate_general_and_int32(tmp_path):
    rng = np.random.default_rng(0)
    A = sp.random(40, 40, density=0.1, random_state=1, format="coo") * 3.0
    scipy.io.mmwrite(str(tmp_path / "gen.mtx"), A)
    B = RIO.loaddata("gen", data_dir=str(tmp_path))
    assert sp.isspmatrix_csc(B) and B.dtype == np.complex128
    assert B.indices.dtype == np.int32 and B.indptr.dtype == np.int32
    assert np.allclose(B.toarray(), A.toarray())
    B64 = RIO.loaddata("gen", data_dir=str(tmp_path), use_int32_index=False)
    assert B64.indices.dtype == np.int64
    assert np.allclose(B64.toarray(), A.toarray())


def test_loaddata_symmetric_complex_and_dense(tmp_path):
    rng = np.random.default_rng(2)
    M = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    H = sp.coo_matrix(np.tril(M + M.conj().T) * (rng.random((12, 12)) < 0.5))
    Hfull = sp.coo_matrix(H.toarray() + np.tril(H.toarray(), -1).conj().T)
    scipy.io.mmwrite(str(tmp_path / "herm.mtx"), Hfull, symmetry="hermitian")
    B = RIO.loaddata("herm", data_dir=str(tmp_path))
    assert np.allclose(B.toarray(), Hfull.toarray())
    D = rng.standard_normal((5, 7))
    scipy.io.mmwrite(str(tmp_path / "dense.mtx"), D)
    Dd = RIO.loaddata("dense", data_dir=str(tmp_path))
    assert isinstance(Dd, np.ndarray) and Dd.dtype == np.complex128 and np.allclose(Dd, D)


def test_loaddata_missing_file(tmp_path):
    with pytest.raises(FileNotFoundError, match="not found"):
        RIO.loaddata("nope", data_dir=str(tmp_path))


def test_save_grid_result_natural_order(tmp_path):
    rec = R.recursion_grid(-1.0, 1.0, 0.0, 0.5, 5, 4, 2)
    sigma = np.abs(rec.Z) + 0.1          # any function of the point
    RIO.save_grid_result(tmp_path / "out.npz", rec, sigma, mv=np.arange(len(rec.Z)))
    d = np.load(tmp_path / "out.npz")
    X, Y = np.meshgrid(d["x"], d["y"])
    assert np.allclose(d["grid"], np.abs(X + 1j * Y) + 0.1)
    assert np.array_equal(d["sigma"], sigma) and len(d["mv"]) == len(rec.Z)
