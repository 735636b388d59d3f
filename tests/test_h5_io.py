"""elPaSo HDF5 input/output (libelpaso_amd_h5.so), CPU part: symbols, the
/Analysis block of the reference's own input fixture, and the solution and
system-matrix writers against the layout of a solution file elPaSo wrote
(tests/golden/egenoutput_dsg4_excerpt.hdf5, egenoutput_dsg4_freqstep_attrs.txt;
made by tests/golden/make_h5_fixtures.sh)."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from elpaso_amd import h5
from elpaso_amd import ElpError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
FLUID8_2ELEM = os.path.join(GOLD, "acousticmedium_fluid8_2elem_AF_LIN_UAF_ISO_DIR.hdf5")
DSG4_EXCERPT = os.path.join(GOLD, "egenoutput_dsg4_excerpt.hdf5")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

needs_h5dump = pytest.mark.skipif(H5DUMP is None, reason="h5dump (HDF5 tools) not installed")


def _dump(*args):
    return subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True).stdout


def _type_block(text, name):
    """DATATYPE ... block of the dataset or attribute `name` in h5dump -H output."""
    m = re.search(r'(?:DATASET|ATTRIBUTE) "%s" \{\s*(DATATYPE\s+.*?)\s*DATASPACE' % re.escape(name), text, re.S)
    assert m, name
    return re.sub(r"\s+", " ", m.group(1))


def test_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "elpaso_amd_h5.h")).read()
    declared = sorted(set(re.findall(r"\b(elp_h5_\w+)\s*\(", hdr)))
    assert declared == sorted(h5.EXPORTED_SYMBOLS)
    L = h5.load_library_h5()
    for s in declared:
        assert hasattr(L, s), s


def test_analysis_of_the_reference_fixture():
    # /Analysis of the reference's 2-element Fluid8 input: frequency-basic,
    # start 10, delta 10, steps 10, revision 6, hdf5 1
    a = h5.read_analysis(FLUID8_2ELEM)
    assert a.type == "frequency-basic"
    assert a.description == "my problem"
    assert (a.start, a.delta, a.steps, a.revision, a.write_hdf5) == (10.0, 10.0, 10, 6, True)
    np.testing.assert_array_equal(a.frequencies(), 10.0 + 10.0 * np.arange(10))


@needs_h5dump
def test_equivalent_fluid_fixture_differs_only_in_its_material():
    """The reference's equivalent-fluid fixture: same analysis as the ideal
    fluid one, material type AF_LIN_EQF_ISO_DIR as the HDF5 tools read it
    (the GPU test checks that an element group using it is rejected)."""
    eqf = os.path.join(GOLD, "acousticmedium_fluid8_2elem_AF_LIN_EQF_ISO_DIR.hdf5")
    assert h5.read_analysis(eqf) == h5.read_analysis(FLUID8_2ELEM)
    attrs = _dump("-A", "-d", "/Materials/material1", eqf)
    assert re.search(r'ATTRIBUTE "MaterialType".*?\(0\): "AF_LIN_EQF_ISO_DIR"', attrs, re.S)
    for name in ("/Nodes/mtxFemNodes", "/Elements/mtxFemElemGroup1"):
        assert _dump("-d", name, eqf).split("\n", 1)[1] == _dump("-d", name, FLUID8_2ELEM).split("\n", 1)[1]


def test_reads_a_solution_elpaso_wrote():
    x1 = h5.read_complex_vector(DSG4_EXCERPT, "/Solution/State/vecFemStep1", 210)
    x2 = h5.read_complex_vector(DSG4_EXCERPT, "/Solution/State/vecFemStep2", 210)
    assert np.all(np.isfinite(x1)) and np.all(np.isfinite(x2))
    assert np.abs(x1).max() > 0 and not np.array_equal(x1, x2)
    if H5DUMP:  # the same values as the HDF5 tools print them
        txt = _dump("-m", "%.17g", "-d", "/Solution/State/vecFemStep1", DSG4_EXCERPT)
        body = txt.split("DATA {", 1)[1]
        nums = [float(t) for t in re.findall(r"(?<![\w(])-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?(?![\w)])", body)]
        ref = np.array(nums).reshape(-1, 2)
        np.testing.assert_array_equal(x1.real, ref[:, 0])
        np.testing.assert_array_equal(x1.imag, ref[:, 1])
    with pytest.raises(ElpError, match="length"):
        h5.read_complex_vector(DSG4_EXCERPT, "/Solution/State/vecFemStep1", 209)
    with pytest.raises(ElpError, match="missing dataset"):
        h5.read_complex_vector(DSG4_EXCERPT, "/Solution/State/vecFemStep9", 210)


def _solution(tmp_path, n=7, nf=3, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((nf, n)) + 1j * rng.standard_normal((nf, n))
    freqs = 10.0 + 1.0 * np.arange(nf)
    ids = np.repeat(np.arange(1, (n + 2) // 3 + 1), 3)[:n].astype(np.int64)
    dofs = np.tile([0, 1, 2], n)[:n].astype(np.int32)
    path = tmp_path / "sol.hdf5"
    h5.write_solution(path, freqs, X, ids, dofs)
    return path, X, freqs, ids, dofs


def test_solution_round_trip_bit_exact(tmp_path):
    path, X, freqs, ids, dofs = _solution(tmp_path)
    for k in range(len(freqs)):
        got = h5.read_complex_vector(path, f"/Solution/State/vecFemStep{k + 1}", X.shape[1])
        np.testing.assert_array_equal(got, X[k])


@needs_h5dump
def test_solution_layout_matches_elpaso_output(tmp_path):
    path, X, freqs, ids, dofs = _solution(tmp_path)
    ours = _dump("-H", "-A", str(path))
    ref = _dump("-H", DSG4_EXCERPT)
    assert _type_block(ours, "vecFemStep1") == _type_block(ref, "vecFemStep1")
    assert _type_block(ours, "mtxDofMap") == _type_block(ref, "mtxDofMap")
    ref_attr = open(os.path.join(GOLD, "egenoutput_dsg4_freqstep_attrs.txt")).read()
    assert _type_block(ours, "FreqStep1") == _type_block(ref_attr, "FreqStep1")
    # attribute text: std::to_string of the frequency ("10.000000", "11.000000")
    vals = _dump("-a", "/Solution/State/FreqStep1", "-a", "/Solution/State/FreqStep2", str(path))
    ref_vals = re.findall(r'\(0\): "([^"]*)"', ref_attr)
    assert re.findall(r'\(0\): "([^"]*)"', vals) == ref_vals == ["10.000000", "11.000000"]
    m = _dump("-d", "/Solution/Maps/mtxDofMap", str(path))
    nums = [int(t) for t in re.findall(r"(?<![\w(])-?\d+(?![\w)])", m.split("DATA {", 1)[1])]
    np.testing.assert_array_equal(np.array(nums).reshape(-1, 2), np.stack([ids, dofs], axis=1))


def test_solution_file_is_truncated(tmp_path):
    path, _, _, _, _ = _solution(tmp_path, nf=3)
    path2, X, _, _, _ = _solution(tmp_path, nf=1, seed=2)
    assert path == path2
    with pytest.raises(ElpError, match="missing dataset"):
        h5.read_complex_vector(path, "/Solution/State/vecFemStep2", X.shape[1])


@needs_h5dump
def test_csr_export_layout(tmp_path):
    # A 3x3 system in 0-based CSR; the file holds it 1-based (mathlibrary.cpp:62)
    rp = np.array([0, 2, 3, 5], np.int32)
    ci = np.array([0, 2, 1, 0, 2], np.int32)
    v = np.array([1 + 2j, -3, 4j, 5.5, 6 - 1j])
    path = tmp_path / "sys.hdf5"
    h5.write_csr(path, "K", rp, ci, v)
    h5.write_csr(path, "M", rp, ci, 2 * v)  # a second matrix in the same file
    np.testing.assert_array_equal(h5.read_complex_vector(path, "/SystemMatrices/K/cmpCsrVal", 5), v)
    np.testing.assert_array_equal(h5.read_complex_vector(path, "/SystemMatrices/M/cmpCsrVal", 5), 2 * v)
    txt = _dump("-A", "-g", "/SystemMatrices/K", str(path))
    assert _type_block(txt, "vecCsrIa") == "DATATYPE H5T_STD_I32LE"
    assert _type_block(txt, "vecCsrJa") == "DATATYPE H5T_STD_I32LE"
    assert _type_block(txt, "numrow") == "DATATYPE H5T_STD_I32LE"
    ia = _dump("-d", "/SystemMatrices/K/vecCsrIa", str(path)).split("DATA {", 1)[1]
    ja = _dump("-d", "/SystemMatrices/K/vecCsrJa", str(path)).split("DATA {", 1)[1]
    assert [int(t) for t in re.findall(r"(?<![\w(])\d+(?![\w)])", ia)] == [1, 3, 4, 6]
    assert [int(t) for t in re.findall(r"(?<![\w(])\d+(?![\w)])", ja)] == [1, 3, 2, 1, 3]
    with pytest.raises(ElpError, match="exists already"):
        h5.write_csr(path, "K", rp, ci, v)
    # a malformed pattern is an error, and nothing of it is written
    with pytest.raises(ElpError, match="column index out of range at entry 4"):
        h5.write_csr(path, "C1", rp, np.array([0, 2, 1, 0, 3], np.int32), v)
    with pytest.raises(ElpError, match="rowptr decreases at row 1"):
        h5.write_csr(path, "C2", np.array([0, 3, 2, 5], np.int32), ci, v)
    assert "C1" not in _dump("-n", str(path)) and "C2" not in _dump("-n", str(path))


def test_errors_are_reported_not_printed(tmp_path):
    path, _, _, _, _ = _solution(tmp_path)
    with pytest.raises(ElpError, match="missing group /Analysis"):
        h5.read_analysis(path)
    with pytest.raises(ElpError, match="cannot open"):
        h5.read_analysis(tmp_path / "absent.hdf5")
    lib = h5.load_library_h5()
    assert lib.elp_h5_read_analysis(None, None) != 0
