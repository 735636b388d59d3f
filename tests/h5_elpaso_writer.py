"""Test helper: writes a Model as an elPaSo HDF5 input file, with the layout
the reference's parser reads (misc/hdf5/rulesh5.h; types as in the
reference's own input fixture acousticmedium_fluid8_2elem_AF_LIN_UAF_ISO_DIR.hdf5:
u64 element rows [id, nodes...], compound nodes {Ids i64, xCoords, yCoords,
zCoords f64}, material parameters as strings, constraint flags u8, values
f64, interface rows i64 [id, ori, structure nodes, fluid nodes]).

Calls libhdf5 1.10 through ctypes (h5py is not installed).  Only the tests
use it: the library reads these files, it does not write inputs."""
from __future__ import annotations

import ctypes as C

import numpy as np

H5LIB = "/opt/conda/lib/libhdf5.so.103"
_L = None


def _lib():
    global _L
    if _L is None:
        L = C.CDLL(H5LIB)
        L.H5open()
        hid, sz = C.c_int64, C.c_size_t
        L.H5Fcreate.restype = hid
        L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid, hid]
        L.H5Gcreate2.restype = hid
        L.H5Gcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid]
        L.H5Screate.restype = hid
        L.H5Screate.argtypes = [C.c_int]
        L.H5Screate_simple.restype = hid
        L.H5Screate_simple.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.H5Dcreate2.restype = hid
        L.H5Dcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid, hid]
        L.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
        L.H5Acreate2.restype = hid
        L.H5Acreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid]
        L.H5Awrite.argtypes = [hid, hid, C.c_void_p]
        L.H5Tcopy.restype = hid
        L.H5Tcopy.argtypes = [hid]
        L.H5Tset_size.argtypes = [hid, sz]
        L.H5Tset_cset.argtypes = [hid, C.c_int]
        L.H5Tcreate.restype = hid
        L.H5Tcreate.argtypes = [C.c_int, sz]
        L.H5Tinsert.argtypes = [hid, C.c_char_p, sz, hid]
        for f in ("H5Fclose", "H5Gclose", "H5Sclose", "H5Dclose", "H5Aclose", "H5Tclose"):
            getattr(L, f).argtypes = [hid]
        _L = L
    return _L


def _t(name):
    return C.c_int64.in_dll(_lib(), name).value


H5P_DEFAULT, H5F_ACC_TRUNC, H5S_SCALAR, H5S_ALL = 0, 2, 0, 0
H5T_COMPOUND, H5T_CSET_UTF8 = 6, 1
H5T_VARIABLE = C.c_size_t(-1).value


class _File:
    def __init__(self, path):
        self.L = _lib()
        self.f = self.L.H5Fcreate(str(path).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        assert self.f >= 0, path
        self.open = []

    def group(self, name):
        g = self.L.H5Gcreate2(self.f, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        assert g >= 0, name
        self.open.append(("H5Gclose", g))
        return g

    def dataset(self, loc, name, file_type, mem_type, arr, dims=None):
        dims = arr.shape if dims is None else dims
        d = (C.c_uint64 * len(dims))(*dims)
        sp = self.L.H5Screate_simple(len(dims), d, None)
        ds = self.L.H5Dcreate2(loc, name.encode(), file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        assert ds >= 0, name
        if arr.size:
            assert self.L.H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data) >= 0
        self.L.H5Sclose(sp)
        self.open.append(("H5Dclose", ds))
        return ds

    def empty(self, loc, name):
        return self.dataset(loc, name, _t("H5T_IEEE_F64LE_g"), _t("H5T_NATIVE_DOUBLE_g"), np.zeros(0), dims=(0,))

    def attr(self, obj, name, value):
        L = self.L
        sp = L.H5Screate(H5S_SCALAR)
        if isinstance(value, str):
            t = L.H5Tcopy(_t("H5T_C_S1_g"))
            L.H5Tset_size(t, H5T_VARIABLE)
            L.H5Tset_cset(t, H5T_CSET_UTF8)
            buf = C.c_char_p(value.encode())
            a = L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT)
            assert L.H5Awrite(a, t, C.byref(buf)) >= 0
            L.H5Tclose(t)
        else:
            ftype, mtype, cval = {
                "u8": ("H5T_STD_U8LE_g", "H5T_NATIVE_UCHAR_g", C.c_uint8),
                "u64": ("H5T_STD_U64LE_g", "H5T_NATIVE_ULLONG_g", C.c_uint64),
                "i64": ("H5T_STD_I64LE_g", "H5T_NATIVE_LLONG_g", C.c_int64),
                "f64": ("H5T_IEEE_F64LE_g", "H5T_NATIVE_DOUBLE_g", C.c_double),
            }[value[0]]
            v = cval(value[1])
            a = L.H5Acreate2(obj, name.encode(), _t(ftype), sp, H5P_DEFAULT, H5P_DEFAULT)
            assert L.H5Awrite(a, _t(mtype), C.byref(v)) >= 0
        L.H5Aclose(a)
        L.H5Sclose(sp)

    def close(self):
        for f, h in reversed(self.open):
            getattr(self.L, f)(h)
        self.L.H5Fclose(self.f)


_KIND_NAMES = {1: "Fluid8", 2: "Brick8", 3: "Fluid20", 4: "Brick20"}


def write_input(path, model, start=10.0, delta=10.0, steps=10, revision=6, element_type=None,
                material_type=None, elem_load=False, interface_type="InterfaceMindlin", valp=None,
                nodes_rank2=False):
    """Writes model (elpaso_amd.Model) as an elPaSo input file.  element_type
    / material_type replace the first group's ElementType / the first
    material's MaterialType, elem_load adds a "vn" element load, and
    interface_type names the interface elements' type (content the library
    must reject).  An acoustic constraint's value goes to attribute "p"
    (u8 when it is a small whole number, as in the reference's fixture, f64
    otherwise) and to "valp", unless valp gives another "valp" value.
    nodes_rank2 writes the node records as an n x 1 dataset (malformed: the
    reader expects a 1-D dataset)."""
    F = _File(path)
    g = F.group("/Analysis")
    F.attr(g, "type", "frequency-basic")
    F.attr(g, "description", model.name)
    F.attr(g, "start", ("f64", start))
    F.attr(g, "delta", ("f64", delta))
    F.attr(g, "steps", ("u64", steps))
    F.attr(g, "revision", ("u64", revision))
    F.attr(g, "hdf5", ("u8", 1))
    # nodes: compound {Ids i64, xCoords, yCoords, zCoords f64}
    g = F.group("/Nodes")
    rec = np.zeros(len(model.node_ids), dtype=[("Ids", "<i8"), ("xCoords", "<f8"), ("yCoords", "<f8"),
                                               ("zCoords", "<f8")])
    rec["Ids"] = model.node_ids
    rec["xCoords"], rec["yCoords"], rec["zCoords"] = model.xyz[:, 0], model.xyz[:, 1], model.xyz[:, 2]
    L = F.L
    ct = L.H5Tcreate(H5T_COMPOUND, 32)
    L.H5Tinsert(ct, b"Ids", 0, _t("H5T_STD_I64LE_g"))
    for k, nm in enumerate((b"xCoords", b"yCoords", b"zCoords")):
        L.H5Tinsert(ct, nm, 8 + 8 * k, _t("H5T_IEEE_F64LE_g"))
    F.dataset(g, "mtxFemNodes", ct, ct, rec, dims=(len(rec), 1) if nodes_rank2 else None)
    L.H5Tclose(ct)
    # materials: parameters as strings
    g = F.group("/Materials")
    F.attr(g, "N", ("i64", len(model.materials)))
    for mid, m in model.materials.items():
        d = F.empty(g, f"material{mid}")
        F.attr(d, "Id", ("u64", mid))
        F.attr(d, "Name", "name")
        if material_type is not None and mid == next(iter(model.materials)):
            F.attr(d, "MaterialType", material_type)
        elif m.kind == "fluid":
            F.attr(d, "MaterialType", "AF_LIN_UAF_ISO_DIR")
            F.attr(d, "c", repr(float(m.c)))
            F.attr(d, "rho", repr(float(m.rho)))
            F.attr(d, "t", "0.0")
        elif m.eta:
            F.attr(d, "MaterialType", "STR_LIN_VIS_ISO_DIR")
            for k, v in (("E", m.E), ("nu", m.nu), ("rho", m.rho), ("eta", m.eta)):
                F.attr(d, k, repr(float(v)))
            F.attr(d, "type", str(int(m.visco_type)))
        else:
            F.attr(d, "MaterialType", "STR_LIN_ELA_ISO_DIR")
            for k, v in (("E", m.E), ("nu", m.nu), ("rho", m.rho)):
                F.attr(d, k, repr(float(v)))
    # element groups: u64 rows [element id, node ids...]
    g = F.group("/Elements")
    for k, eg in enumerate(model.groups):
        rows = np.concatenate([np.asarray(eg.elem_ids, np.uint64)[:, None], np.asarray(eg.conn, np.uint64)], axis=1)
        rows = np.ascontiguousarray(rows)
        d = F.dataset(g, f"mtxFemElemGroup{k + 1}", _t("H5T_STD_U64LE_g"), _t("H5T_NATIVE_ULLONG_g"), rows)
        F.attr(d, "ElementType", element_type if (element_type and k == 0) else _KIND_NAMES[eg.kind])
        F.attr(d, "Id", ("u64", k + 1))
        F.attr(d, "MaterialId", ("i64", eg.material_id))
        F.attr(d, "N", ("u64", len(eg.elem_ids)))
        F.attr(d, "Name", f"Block_{k + 1}")
    g = F.group("/ElemLoads")
    if elem_load:
        d = F.empty(g, "elemLoad1")
        F.attr(d, "LoadType", "vn")
        F.attr(d, "Id", ("u64", 1))
        F.attr(d, "ElementId", ("u64", int(model.groups[0].elem_ids[0])))
        F.attr(d, "Face", ("u64", 1))
        F.attr(d, "vn", ("f64", 1.0))
    # constraints: one dataset per (node, dof) with the flag and value of that dof
    flags = ["u1", "u2", "u3", "w1", "w2", "w3", "p0", "p1", "p2", "p3", "xd3", "wd1", "wd2", "dwdx", "dwdy",
             "dwdxy", "fluid", "z_1", "z_3", "x1_2", "x2_2"]
    g = F.group("/NodeConstraints")
    for k, (nid, dof, v) in enumerate(model.dirichlet):
        d = F.empty(g, f"nodeConstraint{k + 1}_{nid}")
        F.attr(d, "Id", ("u64", k + 1))
        F.attr(d, "NodeId", ("u64", nid))
        if dof == 16:
            F.attr(d, "ConstraintType", "acoustic")
            pv = complex(v).real
            F.attr(d, "p", ("u8", int(pv)) if pv == int(pv) and 0 <= pv < 256 else ("f64", pv))
            F.attr(d, "valp", ("f64", pv if valp is None else float(valp)))
        else:
            F.attr(d, "ConstraintType", "structure")
            for j, fl in enumerate(flags):
                F.attr(d, fl, ("u8", int(j == dof)))
                F.attr(d, "val" + fl, ("f64", complex(v).real if j == dof else 0.0))
    g = F.group("/NodeLoads")
    k = 0
    for nid, fx, fy, fz in model.forces:
        k += 1
        d = F.empty(g, f"nodeLoad{k}")
        F.attr(d, "LoadType", "point_force")
        F.attr(d, "Id", ("u64", k))
        F.attr(d, "NodeId", ("u64", nid))
        for nm, v in (("Fx", fx), ("Fy", fy), ("Fz", fz)):
            F.attr(d, nm, ("f64", v))
    for nid, Q in model.sources:
        k += 1
        d = F.empty(g, f"nodeLoad{k}")
        F.attr(d, "LoadType", "fluid")
        F.attr(d, "Id", ("u64", k))
        F.attr(d, "NodeId", ("u64", nid))
        F.attr(d, "P", ("f64", Q))
    g = F.group("/InterfaceElements")
    n_if = 0 if model.if_ids is None else len(model.if_ids)
    F.attr(g, "N", ("u64", n_if))
    if n_if:
        rows = np.concatenate([np.asarray(model.if_ids, np.int64)[:, None],
                               np.asarray(model.if_ori, np.int64)[:, None],
                               np.asarray(model.if_struct_nodes, np.int64),
                               np.asarray(model.if_fluid_nodes, np.int64)], axis=1)
        rows = np.ascontiguousarray(rows)
        d = F.dataset(g, "InterfaceElements0", _t("H5T_STD_I64LE_g"), _t("H5T_NATIVE_LLONG_g"), rows)
        F.attr(d, "type", interface_type)
        F.attr(d, "NfluidNodes", ("u64", 4))
        F.attr(d, "NstructNodes", ("u64", 4))
        F.attr(d, "matF", ("i64", 1))
        F.attr(d, "mats", ("i64", 2))
    g = F.group("/NCInterfaceElements")
    F.attr(g, "N", ("u64", 0))
    F.close()
