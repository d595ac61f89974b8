"""MAT-file I/O of the C-ABI (ksvd_mat_read / ksvd_mat_write, host only) and
the ksvd_main driver's argument handling (no device work).

Reader pinned by the reference's own MATLAB-written files (tests/golden/
ref_D.mat, ref_temp.mat, kept byte for byte, miCOMPRESSED) against scipy's
independent reader, and by scipy-written files of several storage types;
writer pinned by reading its files back with scipy."""
import json
import os
import subprocess

import numpy as np
import pytest
import scipy.io as sio

from ksvd_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MAIN = os.path.join(os.path.dirname(HERE), "ksvd_amd", "ksvd_main")


def test_reads_reference_dictionary_file():
    D = _lib.mat_read(os.path.join(GOLD, "ref_D.mat"), "D")
    ref = np.load(os.path.join(GOLD, "ref_D.npy"))
    assert D.shape == (300, 9)
    np.testing.assert_array_equal(D, ref)


def test_reads_reference_gamma_file():
    G = _lib.mat_read(os.path.join(GOLD, "ref_temp.mat"), "CUDA_Gamma")
    st = json.load(open(os.path.join(GOLD, "ref_gamma_structure.json")))["K-SVD/K-SVD/output/temp.mat"]
    assert list(G.shape) == st["shape"]
    nz = G != 0
    assert sorted(set(nz.sum(0).tolist())) == st["col_nnz"]
    assert nz.sum(1).tolist() == st["row_nnz"]
    assert np.abs(G).max() == st["abs_max"]
    assert np.abs(G[nz]).min() == st["abs_min_nonzero"]
    np.testing.assert_allclose(G.sum(), st["sum"], rtol=1e-12)
    np.testing.assert_allclose(np.abs(G).sum(), st["abs_sum"], rtol=1e-12)


@pytest.mark.parametrize("compress", [False, True])
def test_reads_scipy_written_files(tmp_path, compress):
    rng = np.random.default_rng(5)
    arrays = {
        "X": rng.standard_normal((37, 53)),
        "S": rng.standard_normal((9, 4)).astype(np.float32),
        "I": rng.integers(-1000, 1000, (5, 7)).astype(np.int32),
        "U": rng.integers(0, 255, (3, 11)).astype(np.uint8),
        "one": np.array([[2.5]]),
        "longname_variable": rng.standard_normal((2, 3)),
    }
    p = str(tmp_path / "in.mat")
    sio.savemat(p, arrays, do_compression=compress)
    for k, a in arrays.items():
        np.testing.assert_array_equal(_lib.mat_read(p, k), a.astype(np.float64))


def test_writer_round_trips_through_scipy(tmp_path):
    rng = np.random.default_rng(6)
    X = rng.standard_normal((64, 100))
    G = np.zeros((9, 3600))
    G[rng.integers(0, 9, 3600), np.arange(3600)] = rng.standard_normal(3600)
    p = str(tmp_path / "out.mat")
    _lib.mat_write(p, {"X": X, "CUDA_Gamma": G, "e": np.zeros((0, 3))}, global_vars=True)
    back = sio.loadmat(p)
    np.testing.assert_array_equal(back["X"], X)
    np.testing.assert_array_equal(back["CUDA_Gamma"], G)
    assert back["e"].shape == (0, 3)
    np.testing.assert_array_equal(_lib.mat_read(p, "CUDA_Gamma"), G)
    np.testing.assert_array_equal(_lib.mat_read(p, "X"), X)


def test_reader_errors(tmp_path):
    p = str(tmp_path / "c.mat")
    sio.savemat(p, {"Z": np.ones((2, 2)) * (1 + 2j), "V": np.ones((2, 2, 2)), "A": np.ones((3, 3))})
    for name, msg in [("Z", "complex"), ("V", "2-D"), ("missing", "no variable")]:
        with pytest.raises(_lib.KsvdError, match=msg):
            _lib.mat_read(p, name)
    bad = tmp_path / "bad.mat"
    bad.write_bytes(b"not a mat file" * 20)
    with pytest.raises(_lib.KsvdError, match="level-5"):
        _lib.mat_read(str(bad), "A")
    with pytest.raises(_lib.KsvdError, match="cannot open"):
        _lib.mat_read(str(tmp_path / "absent.mat"), "A")


def test_reader_rejects_negative_dimensions(tmp_path):
    """A crafted file whose dimensions are -1 x -8 over 8 doubles: the product
    of the two negative sizes equals the element count, so only an explicit
    sign check stops it."""
    p = str(tmp_path / "neg.mat")
    sio.savemat(p, {"A": np.arange(8.0).reshape(1, 8)}, do_compression=False)
    raw = bytearray(open(p, "rb").read())
    # 128-byte header, miMATRIX tag (8), array-flags subelement (16), dims tag (8)
    off = 128 + 8 + 16 + 8
    assert np.frombuffer(bytes(raw[off:off + 8]), "<i4").tolist() == [1, 8]
    raw[off:off + 8] = np.array([-1, -8], "<i4").tobytes()
    open(p, "wb").write(bytes(raw))
    with pytest.raises(_lib.KsvdError, match="negative dimension"):
        _lib.mat_read(p, "A")


def test_driver_arguments():
    assert os.path.exists(MAIN), "ksvd_main not built"
    r = subprocess.run([MAIN, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "CUDA_Gamma" in r.stdout
    r = subprocess.run([MAIN, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown argument" in r.stderr
    r = subprocess.run([MAIN, "--X", "/nonexistent/X.mat"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "Error reading X" in r.stderr
