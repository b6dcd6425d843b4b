"""The C++ programs built against the drop-in header (hlanc_amd/demo/): the
reference demo's four scenarios, the solver CLI and its residual verifier (GPU),
and the format converter (CPU)."""
import os
import subprocess

import numpy as np
import pytest

import refformats as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hlanc_amd", "demo", "_build")


def prog(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} not built (run __graft_entry__.build())")
    return p


def test_matrix_convertor_writes_reference_bin_layout(tmp_path):
    out = tmp_path / "j.bin"
    r = subprocess.run([prog("matrix_convertor"), "mm", os.path.join(R.GOLDEN, "Journals.mtx"), str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.int32)
    ncols, nrows = raw[0], raw[1]
    rp = raw[2:2 + nrows + 1]
    nnz = raw[2 + nrows + 1]
    assert (ncols, nrows, nnz) == (124, 124, 124 + 2 * 5972) and rp[-1] == nnz
    col = raw[3 + nrows + 1:3 + nrows + 1 + nnz]
    val = np.frombuffer(out.read_bytes()[4 * (3 + nrows + 1 + nnz):], dtype=np.float64)
    A = R.to_dense(rp, col, val, 124, False)
    J = R.to_dense(*R.journals())
    np.testing.assert_array_equal(A, J)


@pytest.mark.gpu
def test_demo_scenarios(tmp_path):
    r = subprocess.run([prog("hlanc_demo"), R.GOLDEN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stderr)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_solver_cli_and_verifier(tmp_path):
    out = tmp_path / "eigs.txt"
    with open(out, "w") as f:
        r = subprocess.run([prog("hlanc_solver_gpu"), "mm", os.path.join(R.GOLDEN, "Journals.mtx"), "4", "12"],
                           stdout=f, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().split("\n")
    assert lines[0] == "4" and lines[1] == "124"
    vals = np.array([float(x) for x in lines[2:6]])
    np.testing.assert_allclose(vals, [19795.37857462678, 23432.29299233234, 25008.48027847430, 68542.71718825979],
                               rtol=1e-13)
    v = subprocess.run([prog("verify_eigs"), "mm", os.path.join(R.GOLDEN, "Journals.mtx"), str(out), "0", "1e-12"],
                       capture_output=True, text=True, timeout=300)
    assert v.returncode == 0, v.stderr
    assert v.stderr.count("max r-diff") == 4


@pytest.mark.gpu
def test_batch_solver_cli_is_bitwise_the_single_solves():
    j = os.path.join(R.GOLDEN, "Journals.mtx")
    s = os.path.join(R.GOLDEN, "simple.pp")
    r = subprocess.run([prog("hlanc_batch_solver"), "4", "12", "mm", j, "pp", s, "mm", j],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "0 mismatches" in r.stderr
    lines = r.stdout.split("\n")
    assert lines[0] == "4" and lines[1] == "124"
    np.testing.assert_allclose([float(x) for x in lines[2:6]],
                               [19795.37857462678, 23432.29299233234, 25008.48027847430, 68542.71718825979],
                               rtol=1e-13)
    assert lines[6] == "4" and lines[7] == "16"
    np.testing.assert_allclose([float(x) for x in lines[8:12]],
                               [2.337028429548383, 4.497836710984441, 12.34353751967706, 110.4221400959146],
                               rtol=1e-13)
