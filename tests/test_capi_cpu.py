"""CPU-side checks of the product library: it loads, exports every C-ABI entry
point declared in include/*.h, reads and writes the reference's file formats,
and fails loudly (not silently on the CPU) when no GPU is present."""
import ctypes
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import hlanc_amd as H
import refformats as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(hl_[a-z0-9_]+)\s*\(", text):
            if not m.group(1).endswith("_fn"):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = H.lib()
    declared = declared_symbols()
    assert "hl_ds_solve_csr" in declared and "hl_csr_spmv" in declared
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (hl_[a-z0-9_]+)", out))
    assert declared <= exported, declared - exported
    for name in declared:
        assert hasattr(lib, name)
    assert set(H.exported_symbols()) <= exported


def test_version_matches_reference_macros():
    assert H.lib().hl_version().decode() == "1.0"  # HLANC_VERSION_MAJOR/MINOR (HLanc.hpp:15-16)


def test_no_gpu_means_loud_failure():
    if H.device_count() > 0:
        pytest.skip("a GPU is visible")
    m = H.dcsr_matrix_t(row_ptr=[0, 1], col=[0], val=[1.0])
    with pytest.raises(H.HLancError):
        H.dcsrmv_operator_gpu(m)
    with pytest.raises(H.HLancError):
        H.ds_solver_gpu(16, 4, 12, lambda x, y: None)


@pytest.mark.parametrize("which", ["simple", "journals"])
def test_readers_match_independent_parse(which):
    if which == "simple":
        m = H.dcsr_matrix_t("pp", os.path.join(R.GOLDEN, "simple.pp"))
        rp, col, val, nc, sym = R.simple()
    else:
        m = H.dcsr_matrix_t("mm", os.path.join(R.GOLDEN, "Journals.mtx"))
        rp, col, val, nc, sym = R.journals()
    np.testing.assert_array_equal(m.get_row_ptr(), rp)
    np.testing.assert_array_equal(m.get_column_index(), col)
    np.testing.assert_array_equal(m.get_value(), val)
    assert m.get_num_columns() == nc and m.symmetric() == sym
    assert m.get_nnz() == (256 if which == "simple" else 6096)


def test_pp_symmetric_flag_drops_upper(tmp_path):
    m = H.dcsr_matrix_t("pp", os.path.join(R.GOLDEN, "simple.pp"), True)
    rp, col, val, nc, sym = R.read_pp(os.path.join(R.GOLDEN, "simple.pp"), True)
    np.testing.assert_array_equal(m.get_row_ptr(), rp)
    np.testing.assert_array_equal(m.get_column_index(), col)
    assert m.symmetric() and m.get_nnz() == 16 * 17 // 2


def test_bin_round_trip_expands_symmetric(tmp_path):
    m = H.dcsr_matrix_t("mm", os.path.join(R.GOLDEN, "Journals.mtx"))
    f = str(tmp_path / "j.bin")
    m.save_bin(f)
    b = H.dcsr_matrix_t("bin", f)
    assert not b.symmetric() and b.get_nnz() == 124 + 2 * 5972
    A = R.to_dense(m.get_row_ptr(), m.get_column_index(), m.get_value(), 124, True)
    B = R.to_dense(b.get_row_ptr(), b.get_column_index(), b.get_value(), 124, False)
    np.testing.assert_array_equal(A, B)
    # byte layout: int ncols, int nrows, int row_ptr[nrows+1], int nnz, int col[nnz], double val[nnz]
    raw = open(f, "rb").read()
    ncols, nrows = struct.unpack_from("ii", raw, 0)
    assert (ncols, nrows) == (124, 124)
    nnz = struct.unpack_from("i", raw, 8 + 4 * 125)[0]
    assert nnz == b.get_nnz() and len(raw) == 8 + 4 * 125 + 4 + 12 * nnz


def test_pp_and_mm_writers(tmp_path):
    m = H.dcsr_matrix_t("pp", os.path.join(R.GOLDEN, "simple.pp"))
    m.save_pp(str(tmp_path / "s.pp"))
    assert open(tmp_path / "s.pp").read() == open(os.path.join(R.GOLDEN, "simple.pp")).read().replace(" \n", "\n")
    m.save_mm(str(tmp_path / "s.mtx"))
    t = open(tmp_path / "s.mtx").read().splitlines()
    assert t[0] == "%%MatrixMarket matrix coordinate real general" and t[1] == "16 16 256"
    back = H.dcsr_matrix_t("mm", str(tmp_path / "s.mtx"))
    np.testing.assert_array_equal(back.get_value(), m.get_value())


def test_reader_errors(tmp_path):
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(H.HLancError) as e:
        H.dcsr_matrix_t("mm", str(bad))
    assert "banner" in str(e.value)
    with pytest.raises(H.HLancError):
        H.dcsr_matrix_t("mm", str(tmp_path / "missing.mtx"))
    with pytest.raises(H.HLancError):
        H.dcsr_matrix_t("xx", str(bad))
    bad.write_text("0:1 1:x\n")
    with pytest.raises(H.HLancError):
        H.dcsr_matrix_t("pp", str(bad))


def test_tuning_keys_round_trip_and_refuse_bad_values():
    # process-wide kernel parameters: no GPU needed to set or read them
    for key, value in (("cgs_blocks_max", 1024), ("dgks_skip_exp", 12), ("f32_phase_tol_exp", 7),
                       ("batch_engines", 1)):
        old = H.tune_get(key)
        try:
            H.tune_set(key, value)
            assert H.tune_get(key) == value
        finally:
            H.tune_set(key, old)
    assert H.tune_get("dgks_skip_exp") == 13 and H.tune_get("batch_engines") == 3
    with pytest.raises(H.HLancError):
        H.tune_set("batch_engines", 9)
    with pytest.raises(H.HLancError):
        H.tune_set("dgks_skip_exp", -1)
    with pytest.raises(H.HLancError):
        H.tune_set("no_such_key", 1)
