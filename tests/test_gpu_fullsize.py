"""Assembly at the workloads' full sizes, checked through properties that do
not depend on mesh size (the oracle's element-by-element path is too slow
for millions of elements; at small sizes it is compared entry by entry in
test_gpu_assembly.py):

- exact DOF and nonzero counts of the sparsity pattern;
- symmetry of K and M;
- null spaces: K times the rigid-body modes of the unconstrained steel block
  and K times a constant pressure field vanish;
- totals: sum of M = 3 rho V (structure) or V / (rho c^2) (fluid), sum of the
  Robin face matrix = face area, sum of the FSI coupling = -2 x interface
  area (both coupling blocks, orientation +1, normal +z; the oracle gives the
  same on small meshes);
- Rayleigh damping C = 1e-5 K + 2 M entry by entry.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from elpaso_amd import GpuModel
from elpaso_amd import _lib
from elpaso_amd import meshgen as MG

pytestmark = pytest.mark.gpu


def _csr(g, which):
    rp, ci = g.pattern()
    return sp.csr_matrix((g.values(which), ci, rp), shape=(g.n, g.n))


def _sym_err(A):
    return abs(A - A.T).max() / abs(A).max()


def _null_err(A, v):
    """max |A v| relative to max (|A| |v|)"""
    return np.abs(A @ v).max() / (abs(A) @ np.abs(v)).max()


def test_config1_steel_block_full(ctx):
    m = MG.steel_block()  # 200 x 200 x 4 Brick8, jittered, 1.0 x 1.0 x 0.02 m
    g = GpuModel(ctx, m)
    g.assemble()
    print(f"[fullsize] steel block n={g.n} nnz={g.nnz}", flush=True)
    assert g.n == 606015 and g.nnz == 42260517
    K, M, C = _csr(g, _lib.MAT_K), _csr(g, _lib.MAT_M), _csr(g, _lib.MAT_C)
    assert _sym_err(K) <= 1e-13 and _sym_err(M) <= 1e-13
    ids, dof = g.dof_map()
    order = np.argsort(np.asarray(m.node_ids))
    xyz = np.asarray(m.xyz).reshape(-1, 3)[order[np.searchsorted(np.asarray(m.node_ids)[order], ids)]]
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    modes = [(dof == d).astype(float) for d in range(3)]  # translations
    modes += [np.where(dof == 0, -y, np.where(dof == 1, x, 0.0)),  # rotations about z, x, y
              np.where(dof == 1, -z, np.where(dof == 2, y, 0.0)),
              np.where(dof == 2, -x, np.where(dof == 0, z, 0.0))]
    for v in modes:
        assert _null_err(K, v) <= 1e-10
    rho, V = 7850.0, 1.0 * 1.0 * 0.02
    assert abs(M.sum() - 3 * rho * V) <= 1e-12 * 3 * rho * V
    a, b = m.rayleigh
    ref = a * K + b * M
    assert abs(C - ref).max() <= 1e-14 * (abs(a * K) + abs(b * M)).max()


def test_config2_hex20_cavity_full(ctx):
    m = MG.acoustic_cavity(100, 80, 60, quadratic=True)
    g = GpuModel(ctx, m)
    g.assemble()
    print(f"[fullsize] Hex20 cavity n={g.n} nnz={g.nnz}", flush=True)
    assert g.n == 1976881 and g.nnz == 113205521
    K, M = _csr(g, _lib.MAT_K), _csr(g, _lib.MAT_M)
    assert _sym_err(K) <= 1e-13 and _sym_err(M) <= 1e-13
    assert _null_err(K, np.ones(g.n)) <= 1e-10
    rho, c, V = 1.21, 343.0, 1.0 * 0.8 * 0.6
    assert abs(M.sum() - V / (rho * c * c)) <= 1e-12 * V / (rho * c * c)
    G = g.values(_lib.MAT_G)
    assert abs(G.sum() - 1.0 * 0.8) <= 1e-12


def test_config3_coupled_full(ctx):
    m = MG.coupled_plate_cavity()  # plate 100 x 100 x 2 Brick8 on a 100 x 100 x 120 Fluid8 cavity
    g = GpuModel(ctx, m)
    g.assemble()
    print(f"[fullsize] coupled n={g.n} nnz={g.nnz}", flush=True)
    assert g.n == 1326130 and g.nnz == 38958430
    S = g.values(_lib.MAT_S)
    assert abs(S.sum() + 2.0 * 1.0 * 1.0) <= 1e-12
    K = _csr(g, _lib.MAT_K)
    _, dof = g.dof_map()
    fl = dof == 16
    Kff = K[fl][:, fl]
    assert _null_err(Kff, np.ones(int(fl.sum()))) <= 1e-10
    Kss = K[~fl][:, ~fl]
    assert _sym_err(Kss) <= 1e-13 and _sym_err(Kff) <= 1e-13


def test_config2_sweep_full_independent_residual(ctx):
    """Three frequencies of the full Hex20 cavity sweep on 32 slots; the FP64
    residual of each returned solution is recomputed on the host from
    elPaSo's system A(w), F (elp_dynamic_matrix, the arrays PETSc would get)."""
    m = MG.acoustic_cavity(100, 80, 60, quadratic=True)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([50.0, 1000.0, 1990.0])
    X, rel, it, _ = g.sweep(freqs, block=32)
    print(f"[fullsize] sweep relres {rel} iters {it}", flush=True)
    assert (rel <= 1e-8).all()
    rp, ci = g.pattern()
    for k, f in enumerate(freqs):
        a, F = g.dynamic_matrix(2 * np.pi * f)
        A = sp.csr_matrix((a, ci, rp), shape=(g.n, g.n))
        r = np.linalg.norm(F - A @ X[k]) / np.linalg.norm(F)
        assert r <= 1.5e-8, (f, r, rel[k])
