"""ctypes bindings of the C-ABI declared in include/elpaso_amd.h."""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

from .model import FLUID8, FLUID20, Model

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
_LIB = None

# every entry point the header declares
EXPORTED_SYMBOLS = [
    "elp_ctx_create", "elp_ctx_destroy", "elp_last_error", "elp_default_solve_opts",
    "elp_model_create", "elp_model_destroy", "elp_model_set_nodes", "elp_model_add_fluid_material",
    "elp_model_add_structure_material", "elp_model_add_elements", "elp_model_add_impedance_faces",
    "elp_model_add_interfaces", "elp_model_add_point_source", "elp_model_add_point_force",
    "elp_model_add_dirichlet", "elp_model_set_rayleigh", "elp_model_set_box", "elp_model_get_box", "elp_model_finalize",
    "elp_model_info", "elp_model_dof_map", "elp_assemble", "elp_assemble_timed", "elp_get_pattern",
    "elp_get_values", "elp_get_device_csr", "elp_dynamic_matrix", "elp_element_matrices", "elp_sweep",
    "elp_results_sizes", "elp_node_results", "elp_node_results_noalloc",
    "elp_server_create", "elp_server_destroy", "elp_server_submit", "elp_server_poll", "elp_server_wait",
    "elp_server_stats",
]

MAT_K, MAT_M, MAT_C, MAT_G, MAT_S = 0, 1, 2, 3, 4
ENOCONV = -6
PREC_FP64, PREC_MIXED = 0, 1
PC_JACOBI, PC_MG = 0, 1


# Handles still open, closed in a fixed order (servers, then models, then
# contexts) by an exit handler that runs at the start of interpreter
# shutdown, while every object is intact.  Objects that outlive it (held by a
# stored traceback, a reference cycle or a module global) are then released
# by the interpreter in no particular order; their __del__ must not call into
# the library any more.
_LIVE = {"server": weakref.WeakSet(), "model": weakref.WeakSet(), "ctx": weakref.WeakSet()}
_SHUT_DOWN = False


def _close_all():
    global _SHUT_DOWN
    for kind in ("server", "model", "ctx"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass
    _SHUT_DOWN = True


atexit.register(_close_all)


class ElpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"elpaso_amd error {code}: {msg}")
        self.code = code


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("method", C.c_int), ("precond", C.c_int), ("block", C.c_int),
                ("rtol", C.c_double), ("max_iters", C.c_int), ("inner_rtol", C.c_double),
                ("shift", C.c_double), ("smooth", C.c_int),
                ("coarse_sweeps", C.c_int), ("mg_levels", C.c_int), ("mg_min_nodes", C.c_int),
                ("op_tile_rows", C.c_int), ("graph", C.c_int)]


class SweepStats(C.Structure):
    _fields_ = [("setup_ms", C.c_double), ("solve_ms", C.c_double), ("spmv_calls", C.c_int64),
                ("spmv_ms", C.c_double), ("spmv_bytes", C.c_int64), ("refinements", C.c_int),
                ("unconverged", C.c_int), ("refactors", C.c_int), ("refactor_ms", C.c_double),
                ("op_tile_rows", C.c_int), ("graph_replays", C.c_int64), ("box_tile_rows", C.c_int)]


def library_path():
    return os.path.join(_PKG, "lib", "libelpaso_amd.so")


def load_library():
    """Load the in-tree HIP library.  Raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise ElpError(-5, f"{path} is missing: build it with __graft_entry__.build() or make -C elpaso-core_amd")
    L = C.CDLL(path)
    vp, i64, dbl = C.c_void_p, C.c_int64, C.c_double
    P = C.POINTER
    L.elp_last_error.restype = C.c_char_p
    L.elp_last_error.argtypes = [vp]
    L.elp_ctx_create.argtypes = [P(vp), C.c_int]
    L.elp_ctx_destroy.argtypes = [vp]
    L.elp_ctx_destroy.restype = None
    L.elp_default_solve_opts.argtypes = [P(SolveOpts)]
    L.elp_model_create.argtypes = [vp, P(vp)]
    L.elp_model_destroy.argtypes = [vp]
    L.elp_model_destroy.restype = None
    L.elp_model_set_nodes.argtypes = [vp, i64, vp, vp]
    L.elp_model_add_fluid_material.argtypes = [vp, i64, dbl, dbl]
    L.elp_model_add_structure_material.argtypes = [vp, i64, dbl, dbl, dbl, dbl, C.c_int]
    L.elp_model_add_elements.argtypes = [vp, C.c_int, i64, i64, vp, vp]
    L.elp_model_add_impedance_faces.argtypes = [vp, i64, vp, vp, dbl, dbl, C.c_int]
    L.elp_model_add_interfaces.argtypes = [vp, i64, vp, vp, vp, vp]
    L.elp_model_add_point_source.argtypes = [vp, i64, dbl]
    L.elp_model_add_point_force.argtypes = [vp, i64, dbl, dbl, dbl]
    L.elp_model_add_dirichlet.argtypes = [vp, i64, C.c_int, dbl, dbl]
    L.elp_model_set_rayleigh.argtypes = [vp, dbl, dbl]
    L.elp_model_set_box.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int]
    L.elp_model_get_box.argtypes = [vp, P(C.c_int), P(C.c_int), vp, vp, vp, P(C.c_int), P(C.c_int)]
    L.elp_model_finalize.argtypes = [vp]
    L.elp_model_info.argtypes = [vp, P(i64), P(i64), P(i64)]
    L.elp_model_dof_map.argtypes = [vp, vp, vp]
    L.elp_assemble.argtypes = [vp, P(dbl)]
    L.elp_assemble_timed.argtypes = [vp, C.c_int, P(dbl)]
    L.elp_get_pattern.argtypes = [vp, vp, vp]
    L.elp_get_values.argtypes = [vp, C.c_int, vp]
    L.elp_get_device_csr.argtypes = [vp, C.c_int, P(vp), P(vp), P(vp)]
    L.elp_dynamic_matrix.argtypes = [vp, dbl, dbl, dbl, vp, vp]
    L.elp_element_matrices.argtypes = [vp, C.c_int, i64, vp, dbl, dbl, dbl, vp, vp]
    L.elp_sweep.argtypes = [vp, i64, vp, dbl, dbl, P(SolveOpts), vp, vp, vp, P(SweepStats)]
    L.elp_results_sizes.argtypes = [vp, P(i64), P(i64)]
    L.elp_node_results.argtypes = [vp, i64, vp, vp, vp, vp, P(dbl)]
    L.elp_node_results_noalloc.argtypes = [vp, i64, vp, vp, vp]
    L.elp_server_create.argtypes = [vp, i64, C.c_int, P(vp)]
    L.elp_server_destroy.argtypes = [vp]
    L.elp_server_destroy.restype = None
    L.elp_server_submit.argtypes = [vp, vp, i64, vp, dbl, dbl, P(SolveOpts), P(i64)]
    L.elp_server_poll.argtypes = [vp, i64]
    L.elp_server_wait.argtypes = [vp, i64, vp, vp, vp, P(dbl)]
    L.elp_server_stats.argtypes = [vp, P(i64), P(i64), P(i64)]
    _LIB = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        msg = load_library().elp_last_error(None)
        raise ElpError(rc, msg.decode() if msg else "")


def default_opts(**kw):
    o = SolveOpts()
    _check(load_library().elp_default_solve_opts(C.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Context:
    def __init__(self, device=0):
        L = load_library()
        self._h = C.c_void_p()
        _check(L.elp_ctx_create(C.byref(self._h), device))
        _LIVE["ctx"].add(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            load_library().elp_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _SHUT_DOWN:
            return
        try:
            self.close()
        except Exception:
            pass

    def element_matrices(self, kind, xyz, p0, p1, p2=0.0):
        """Batch element operators (element/elementfem.h:124-159): xyz (E, npe, 3)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        E, npe = xyz.shape[0], xyz.shape[1]
        nd = 1 if kind in (FLUID8, FLUID20) else 3
        Ke = np.zeros((E, nd * npe, nd * npe))
        Me = np.zeros((E, npe, npe))
        _check(load_library().elp_element_matrices(self._h, kind, E, _ptr(xyz), p0, p1, p2, _ptr(Ke), _ptr(Me)))
        return Ke, Me


class GpuModel:
    """A model uploaded to the device (finalized).  Wraps elp_model*."""

    def __init__(self, ctx: Context, model: Model, box=None, finalize=True):
        L = load_library()
        self.ctx = ctx
        self.model = model
        self._h = C.c_void_p()
        _check(L.elp_model_create(ctx.handle, C.byref(self._h)))
        _LIVE["model"].add(self)
        ids = np.ascontiguousarray(model.node_ids, dtype=np.int64)
        xyz = np.ascontiguousarray(model.xyz, dtype=np.float64)
        _check(L.elp_model_set_nodes(self._h, len(ids), _ptr(ids), _ptr(xyz)))
        for mid, m in model.materials.items():
            if m.kind == "fluid":
                _check(L.elp_model_add_fluid_material(self._h, mid, m.rho, m.c))
            else:
                _check(L.elp_model_add_structure_material(self._h, mid, m.E, m.nu, m.rho, m.eta, m.visco_type))
        for g in model.groups:
            eid = np.ascontiguousarray(g.elem_ids, dtype=np.int64)
            conn = np.ascontiguousarray(g.conn, dtype=np.int64)
            _check(L.elp_model_add_elements(self._h, g.kind, g.material_id, len(eid), _ptr(eid), _ptr(conn)))
        if model.imp_elem_ids is not None and len(model.imp_elem_ids):
            e = np.ascontiguousarray(model.imp_elem_ids, dtype=np.int64)
            f = np.ascontiguousarray(model.imp_faces, dtype=np.int32)
            Z = complex(model.impedance)
            _check(L.elp_model_add_impedance_faces(self._h, len(e), _ptr(e), _ptr(f), Z.real, Z.imag,
                                                   int(bool(model.imp_lumped))))
        if model.if_ids is not None and len(model.if_ids):
            ids_ = np.ascontiguousarray(model.if_ids, dtype=np.int64)
            ori = np.ascontiguousarray(model.if_ori, dtype=np.float64)
            fn = np.ascontiguousarray(model.if_fluid_nodes, dtype=np.int64)
            sn = np.ascontiguousarray(model.if_struct_nodes, dtype=np.int64)
            _check(L.elp_model_add_interfaces(self._h, len(ids_), _ptr(ids_), _ptr(ori), _ptr(fn), _ptr(sn)))
        for nid, Q in model.sources:
            _check(L.elp_model_add_point_source(self._h, int(nid), float(Q)))
        for nid, fx, fy, fz in model.forces:
            _check(L.elp_model_add_point_force(self._h, int(nid), fx, fy, fz))
        for nid, dof, v in model.dirichlet:
            v = complex(v)
            _check(L.elp_model_add_dirichlet(self._h, int(nid), int(dof), v.real, v.imag))
        a, b = model.rayleigh
        if a or b:
            _check(L.elp_model_set_rayleigh(self._h, a, b))
        # box: None = the model's own hint (if any), False = no hint (the
        # library reads the box from the mesh when it can)
        box = None if box is False else (box if box is not None else getattr(model, "box", None))
        if box is not None:
            self.set_box(box)
        if finalize:
            self.finalize()

    @classmethod
    def empty(cls, ctx: Context):
        """An unfinalized model with nothing in it (for readers that fill the
        handle through the C-ABI themselves, e.g. the HDF5 input)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.model = None
        self._h = C.c_void_p()
        _check(load_library().elp_model_create(ctx.handle, C.byref(self._h)))
        _LIVE["model"].add(self)
        return self

    def set_box(self, box):
        o = np.ascontiguousarray(box["origin"], dtype=np.float64)
        ln = np.ascontiguousarray(box["length"], dtype=np.float64)
        ne = np.ascontiguousarray(box["nel"], dtype=np.int32)
        _check(load_library().elp_model_set_box(self._h, _ptr(o), _ptr(ln), _ptr(ne), int(box["order"]),
                                                int(box.get("robin_sides", 0))))

    def get_box(self):
        """The box the solver uses: None, or a dict like set_box's with
        "detected" = True when it was read from the mesh."""
        st, det, order, sides = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        o, ln, ne = np.zeros(3), np.zeros(3), np.zeros(3, dtype=np.int32)
        _check(load_library().elp_model_get_box(self._h, C.byref(st), C.byref(det), _ptr(o), _ptr(ln), _ptr(ne),
                                                C.byref(order), C.byref(sides)))
        if not st.value:
            return None
        return dict(origin=tuple(o), length=tuple(ln), nel=tuple(int(v) for v in ne), order=order.value,
                    robin_sides=sides.value, detected=bool(det.value))

    def finalize(self):
        L = load_library()
        _check(L.elp_model_finalize(self._h))
        n, nnz, b = C.c_int64(), C.c_int64(), C.c_int64()
        _check(L.elp_model_info(self._h, C.byref(n), C.byref(nnz), C.byref(b)))
        self.n, self.nnz, self.device_bytes = n.value, nnz.value, b.value

    @property
    def handle(self):
        return self._h

    def release(self):
        """Hand the handle to a new owner (the server); this object forgets it."""
        h = self._h
        self._h = C.c_void_p()
        return h

    def close(self):
        if self._h:
            load_library().elp_model_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _SHUT_DOWN:
            return
        try:
            self.close()
        except Exception:
            pass

    def assemble(self):
        ms = C.c_double()
        _check(load_library().elp_assemble(self._h, C.byref(ms)))
        return ms.value

    def assemble_timed(self, repeats=10):
        ms = C.c_double()
        _check(load_library().elp_assemble_timed(self._h, repeats, C.byref(ms)))
        return ms.value

    def pattern(self):
        rp = np.zeros(self.n + 1, dtype=np.int32)
        ci = np.zeros(self.nnz, dtype=np.int32)
        _check(load_library().elp_get_pattern(self._h, _ptr(rp), _ptr(ci)))
        return rp, ci

    def values(self, which):
        v = np.zeros(self.nnz)
        _check(load_library().elp_get_values(self._h, which, _ptr(v)))
        return v

    def dof_map(self):
        ids = np.zeros(self.n, dtype=np.int64)
        d = np.zeros(self.n, dtype=np.int32)
        _check(load_library().elp_model_dof_map(self._h, _ptr(ids), _ptr(d)))
        return ids, d

    def dynamic_matrix(self, omega, Z=0j):
        a = np.zeros(self.nnz, dtype=np.complex128)
        f = np.zeros(self.n, dtype=np.complex128)
        Z = complex(Z)
        _check(load_library().elp_dynamic_matrix(self._h, omega, Z.real, Z.imag, _ptr(a), _ptr(f)))
        return a, f

    def sweep(self, freqs, Z=0j, opts=None, want_x=True, **kw):
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        nf = len(freqs)
        o = opts if opts is not None else default_opts(**kw)
        X = np.zeros((nf, self.n), dtype=np.complex128) if want_x else None
        rel = np.zeros(nf)
        it = np.zeros(nf, dtype=np.int32)
        st = SweepStats()
        Z = complex(Z)
        rc = load_library().elp_sweep(self._h, nf, _ptr(freqs), Z.real, Z.imag, C.byref(o),
                                      _ptr(X) if X is not None else None, _ptr(rel), _ptr(it), C.byref(st))
        if rc != ENOCONV:  # ELP_ENOCONV: results are filled, st.unconverged counts the misses
            _check(rc)
        return X, rel, it, st


    def node_results(self, X, freqs, cells=True, nodes=True, timing=None):
        """Brick centroid stresses [nf, bricks, 6] and node results [nf,
        nodes, 9] (sigma 0..5, grad p 6..8; nodes in ascending id) of
        solutions X [nf, n] (elp_node_results).  timing: a dict that gets
        the kernels' device time in "kernel_ms"."""
        L = load_library()
        X = np.ascontiguousarray(X, dtype=np.complex128)
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        if X.shape != (len(freqs), self.n):
            raise ValueError("X must be n_freq x n")
        nn, nb = C.c_int64(), C.c_int64()
        _check(L.elp_results_sizes(self._h, C.byref(nn), C.byref(nb)))
        c = np.zeros((len(freqs), nb.value, 6), np.complex128) if cells else None
        nd = np.zeros((len(freqs), nn.value, 9), np.complex128) if nodes else None
        kms = C.c_double()
        _check(L.elp_node_results(self._h, len(freqs), _ptr(freqs), _ptr(X), _ptr(c) if c is not None else None,
                                  _ptr(nd) if nd is not None else None, C.byref(kms)))
        if timing is not None:
            timing["kernel_ms"] = kms.value
        return c, nd


    def node_results_noalloc(self, X_dev, freqs, nodes_dev):
        """elp_node_results_noalloc: X_dev and nodes_dev are device addresses
        (ints) of n_freq x n and n_freq x nodes x 9 complex arrays."""
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        _check(load_library().elp_node_results_noalloc(self._h, len(freqs), _ptr(freqs), C.c_void_p(X_dev),
                                                       C.c_void_p(nodes_dev)))

    def results_sizes(self):
        nn, nb = C.c_int64(), C.c_int64()
        _check(load_library().elp_results_sizes(self._h, C.byref(nn), C.byref(nb)))
        return nn.value, nb.value


class Server:
    """Continuous-batching sweep server (wraps elp_server*).  Requests are
    admitted in arrival order as slots and the HBM budget allow; each runs
    on its own stream, overlapping with the others on the device."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, max_slots: int = 4):
        self.ctx = ctx
        self._h = C.c_void_p()
        self._n = {}
        _check(load_library().elp_server_create(ctx.handle, int(hbm_budget_bytes), int(max_slots),
                                                C.byref(self._h)))
        _LIVE["server"].add(self)

    def submit(self, gmodel: "GpuModel", freqs, Z=0j, opts=None, **kw):
        """Queue a sweep of ``gmodel`` (the server takes the model over)."""
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        o = opts if opts is not None else default_opts(**kw)
        jid = C.c_int64()
        Z = complex(Z)
        _check(load_library().elp_server_submit(self._h, gmodel.handle, len(freqs), _ptr(freqs), Z.real, Z.imag,
                                                C.byref(o), C.byref(jid)))
        gmodel.release()
        self._n[jid.value] = (gmodel.n, len(freqs))
        return jid.value

    def poll(self, job):
        return load_library().elp_server_poll(self._h, int(job)) == 1

    def wait(self, job, want_x=True):
        n, nf = self._n.pop(job)
        X = np.zeros((nf, n), dtype=np.complex128) if want_x else None
        rel = np.zeros(nf)
        it = np.zeros(nf, dtype=np.int32)
        lat = C.c_double()
        rc = load_library().elp_server_wait(self._h, int(job), _ptr(X) if X is not None else None, _ptr(rel),
                                            _ptr(it), C.byref(lat))
        if rc != ENOCONV:  # ELP_ENOCONV: results are filled, relres shows the misses
            _check(rc)
        return X, rel, it, lat.value

    def stats(self):
        a, q, b = C.c_int64(), C.c_int64(), C.c_int64()
        _check(load_library().elp_server_stats(self._h, C.byref(a), C.byref(q), C.byref(b)))
        return {"admitted": a.value, "queued": q.value, "bytes_in_use": b.value}

    def close(self):
        if self._h:
            load_library().elp_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        if _SHUT_DOWN:
            return
        try:
            self.close()
        except Exception:
            pass
