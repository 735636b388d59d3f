"""ctypes bindings of include/elpaso_amd_h5.h: elPaSo HDF5 model input and
result output (``lib/libelpaso_amd_h5.so``, host code linking libhdf5).

Mirrors the reference's parser and writers (paths relative to
elpasoCore/source of elPaSo-Core):
- ``read_model`` <- cFemParserInterface::parseInputFile
  (misc/parser/femparserinterface.cpp:189-663) over cFemParserHDF5
  (misc/parser/femparserhdf5.cpp:62-560)
- ``Analysis.frequencies`` <- cAnalysisFrequencyBasic::generateFrequencies
  (analysis/frequency/analysisfrequencybasic.cpp:36-72)
- ``write_solution`` <- the HDF5 export of cAnalysisFrequencyBasic::FullRun
  (analysisfrequencybasic.cpp:185-214, 243-265)
- ``write_csr`` <- MathLibrary::ExportPetscMatToHDF5Complex (math/mathlibrary.cpp:53-86)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from ._lib import Context, ElpError, GpuModel, _PKG, _ptr, load_library

_LIB = None
HDF5_LIB = "/opt/conda/lib/libhdf5.so.103"  # the libhdf5 1.10 the library is built against

EXPORTED_SYMBOLS = ["elp_h5_last_error", "elp_h5_read_analysis", "elp_h5_read_model", "elp_h5_write_solution",
                    "elp_h5_write_csr", "elp_h5_read_complex_vector"]


class _Analysis(C.Structure):
    _fields_ = [("type", C.c_char * 64), ("description", C.c_char * 256), ("start", C.c_double),
                ("delta", C.c_double), ("steps", C.c_int64), ("revision", C.c_int64), ("write_hdf5", C.c_int32)]


class _Counts(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("materials", C.c_int64), ("element_groups", C.c_int64),
                ("elements", C.c_int64), ("constraints", C.c_int64), ("node_loads", C.c_int64),
                ("interfaces", C.c_int64)]


@dataclass
class Analysis:
    type: str
    description: str
    start: float
    delta: float
    steps: int
    revision: int
    write_hdf5: bool

    def frequencies(self):
        """f_k = start + delta * k, k < steps (analysisfrequencybasic.cpp:36-72)."""
        return self.start + self.delta * np.arange(self.steps, dtype=np.float64)


def library_path():
    return os.path.join(_PKG, "lib", "libelpaso_amd_h5.so")


def load_library_h5():
    """Load the HDF5 library (after the HIP library it calls).  Raises if it
    has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    load_library()
    path = library_path()
    if not os.path.exists(path):
        raise ElpError(-5, f"{path} is missing: build it with __graft_entry__.build() or make -C elpaso-core_amd")
    try:
        L = C.CDLL(path)
    except OSError:
        # the library finds libhdf5 through the link lib/hdf5/libhdf5.so.103
        # that the build makes; where a copy of the tree lost it, load
        # libhdf5 from its install location first, then the library
        if not os.path.exists(HDF5_LIB):
            raise
        C.CDLL(HDF5_LIB, mode=C.RTLD_GLOBAL)
        L = C.CDLL(path)
    vp, i64, cp = C.c_void_p, C.c_int64, C.c_char_p
    L.elp_h5_last_error.restype = cp
    L.elp_h5_last_error.argtypes = []
    L.elp_h5_read_analysis.argtypes = [cp, C.POINTER(_Analysis)]
    L.elp_h5_read_model.argtypes = [vp, cp, C.POINTER(_Analysis), C.POINTER(_Counts)]
    L.elp_h5_write_solution.argtypes = [cp, i64, i64, vp, vp, vp, vp]
    L.elp_h5_write_csr.argtypes = [cp, cp, i64, vp, vp, vp]
    L.elp_h5_read_complex_vector.argtypes = [cp, cp, i64, vp]
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        raise ElpError(rc, load_library_h5().elp_h5_last_error().decode())


def _path(p):
    return os.fsencode(p)


def _analysis(a: _Analysis) -> Analysis:
    return Analysis(a.type.decode(), a.description.decode(), a.start, a.delta, a.steps, a.revision,
                    bool(a.write_hdf5))


def read_analysis(path) -> Analysis:
    a = _Analysis()
    _check(load_library_h5().elp_h5_read_analysis(_path(path), C.byref(a)))
    return _analysis(a)


def read_model(ctx: Context, path, box=None, finalize=True):
    """Model of an elPaSo input file on the device.  Returns (GpuModel,
    Analysis, counts dict).  ``box`` optionally declares the mesh a structured
    box, as for GpuModel; without it the library reads the box from the mesh
    at finalize when the fluid elements tile a uniform lattice."""
    L = load_library_h5()
    g = GpuModel.empty(ctx)
    a, c = _Analysis(), _Counts()
    _check(L.elp_h5_read_model(g.handle, _path(path), C.byref(a), C.byref(c)))
    if box is not None:
        g.set_box(box)
    if finalize:
        g.finalize()
    counts = {name: getattr(c, name) for name, _ in _Counts._fields_}
    return g, _analysis(a), counts


def write_solution(path, freqs, X, node_ids, dofs):
    """Solution file as elPaSo writes it (X: n_freq x n complex)."""
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.complex128)
    if X.ndim != 2 or X.shape[0] != len(freqs):
        raise ValueError("X must be n_freq x n")
    ids = np.ascontiguousarray(node_ids, dtype=np.int64)
    d = np.ascontiguousarray(dofs, dtype=np.int32)
    n = X.shape[1]
    if len(ids) != n or len(d) != n:
        raise ValueError("node_ids and dofs need one entry per row")
    _check(load_library_h5().elp_h5_write_solution(_path(path), n, len(freqs), _ptr(freqs), _ptr(X), _ptr(ids),
                                                   _ptr(d)))


def write_csr(path, name, rowptr, colind, vals):
    """/SystemMatrices/<name> as ExportPetscMatToHDF5Complex writes it."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(colind, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.complex128)
    n = len(rp) - 1
    if len(ci) != rp[-1] or len(v) != rp[-1]:
        raise ValueError("colind and vals need rowptr[-1] entries")
    _check(load_library_h5().elp_h5_write_csr(_path(path), name.encode(), n, _ptr(rp), _ptr(ci), _ptr(v)))


def read_complex_vector(path, dataset, n):
    X = np.zeros(n, dtype=np.complex128)
    _check(load_library_h5().elp_h5_read_complex_vector(_path(path), dataset.encode(), n, _ptr(X)))
    return X
