"""Assembly at the workloads' full sizes: entry by entry against the oracle
(elPaSo's element loop restated in C, a few seconds per matrix at these
sizes) at the 1e-12 bar of test_gpu_assembly.py, and through properties that
do not depend on mesh size:

- exact DOF and nonzero counts of the sparsity pattern;
- symmetry of K and M;
- null spaces: K times the rigid-body modes of the unconstrained steel block
  and K times a constant pressure field vanish;
- totals: sum of M = 3 rho V (structure) or V / (rho c^2) (fluid), sum of the
  Robin face matrix = face area, sum of the FSI coupling = -2 x interface
  area (both coupling blocks, orientation +1, normal +z; the oracle gives the
  same on small meshes);
- Rayleigh damping C = 1e-5 K + 2 M entry by entry;
- node results: exact pressure gradients of a quadratic field on the Hex20
  cavity and constant stresses of a linear displacement on the jittered
  Brick8 block, at full size.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from elpaso_amd import GpuModel
from elpaso_amd import _lib
from elpaso_amd import meshgen as MG
from test_gpu_assembly import _check_close

pytestmark = pytest.mark.gpu


def _oracle_parity(g, m, which):
    """GPU values of `which` against the oracle's assembly of the same model,
    entry by entry at 1e-12 (normwise and relative to each row's largest)."""
    num = O.Numbering(m)
    rp, ci = O.pattern(m, num)
    grp, gci = g.pattern()
    assert np.array_equal(rp, grp) and np.array_equal(ci, gci), "pattern differs"
    del gci
    ref = O.assemble_km(m, num, (rp, ci))
    refs = {_lib.MAT_K: ref.K, _lib.MAT_M: ref.M, _lib.MAT_C: ref.C, _lib.MAT_G: ref.CG, _lib.MAT_S: ref.S}
    for w in which:
        _check_close(f"{m.name} mat{w}", g.values(w), refs[w], rp)
        print(f"[fullsize] {m.name} mat{w} matches the oracle", flush=True)


def test_config1_steel_block_matches_oracle(ctx):
    """The jittered 200 x 200 x 4 steel block: K, M and C = 1e-5 K + 2 M."""
    m = MG.steel_block()
    g = GpuModel(ctx, m)
    g.assemble()
    _oracle_parity(g, m, (_lib.MAT_K, _lib.MAT_M, _lib.MAT_C))


def test_config2_hex20_cavity_matches_oracle(ctx):
    """The 100 x 80 x 60 Fluid20 cavity: K, M and the Robin face matrix."""
    m = MG.acoustic_cavity(100, 80, 60, quadratic=True)
    g = GpuModel(ctx, m)
    g.assemble()
    _oracle_parity(g, m, (_lib.MAT_K, _lib.MAT_M, _lib.MAT_G))


def test_config3_coupled_matches_oracle(ctx):
    """The plate (100 x 100 x 2 Brick8) on the cavity (100 x 100 x 120
    Fluid8): K, M and the FSI coupling."""
    m = MG.coupled_plate_cavity()
    g = GpuModel(ctx, m)
    g.assemble()
    _oracle_parity(g, m, (_lib.MAT_K, _lib.MAT_M, _lib.MAT_S))


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


def test_config2_node_results_full(ctx):
    """Pressure gradients at the 1.98M nodes of the Hex20 cavity for a
    quadratic pressure field, which the straight-sided 20-node elements
    reproduce exactly: every element's gradient at a node equals the field's,
    so each node holds (elements at the node) x grad p (elPaSo sums, it does
    not average, element/fluid/elementfluid3d.cpp:395-404)."""
    m = MG.acoustic_cavity(100, 80, 60, quadratic=True)
    g = GpuModel(ctx, m)
    ids, dofs = g.dof_map()
    assert (dofs == 16).all() and np.array_equal(ids, np.sort(m.node_ids))
    xyz = m.xyz[np.argsort(m.node_ids)]
    x, y, z = xyz.T
    X = (x * x - 2 * y * z + 3 * z + 0.5j * y * y)[None, :]
    tm = {}
    _, nodes = g.node_results(X, [100.0], cells=False, timing=tm)
    print(f"[fullsize] node results of {g.n} nodes: {tm['kernel_ms']:.2f} ms of kernels", flush=True)
    cnt = np.zeros(len(ids), np.int64)
    np.add.at(cnt, np.searchsorted(ids, m.groups[0].conn.ravel()), 1)
    ref = cnt[:, None] * np.stack([2 * x, -2 * z + 1j * y, -2 * y + 3], axis=1)
    assert np.abs(nodes[0, :, 6:] - ref).max() <= 1e-10 * np.abs(ref).max()
    assert not nodes[0, :, :6].any()


def test_config1_brick_stresses_full(ctx):
    """Centroid stresses of the 160,000 jittered Brick8 elements of the steel
    block for a linear displacement field (reproduced exactly by trilinear
    elements, so the strain is the same everywhere): every cell and every
    node holds D eps, components xx, yy, zz, xy, yz, xz."""
    m = MG.steel_block()
    g = GpuModel(ctx, m)
    ids, dofs = g.dof_map()
    order = np.argsort(m.node_ids)
    xyz = m.xyz[order]
    x, y, z = xyz.T
    u = np.stack([1e-3 * x + 4e-4 * y, -2e-3 * y - 6e-4 * z, 0.5e-3 * z + 9e-4 * x], axis=1)
    rank = np.searchsorted(np.sort(m.node_ids), ids)
    X = (u[rank, dofs] * (1 - 0.25j))[None, :]
    cells, nodes = g.node_results(X, [200.0])
    E, nu = 210e9, 0.3
    f = E / ((1 + nu) * (1 - 2 * nu))
    D = np.zeros((6, 6))
    D[:3, :3] = f * nu
    np.fill_diagonal(D[:3, :3], f * (1 - nu))
    D[3, 3] = D[4, 4] = D[5, 5] = f * (1 - 2 * nu) / 2
    sig = D @ np.array([1e-3, -2e-3, 0.5e-3, 4e-4, -6e-4, 9e-4]) * (1 - 0.25j)
    scale = np.abs(sig).max()
    print(f"[fullsize] brick stresses of {cells.shape[1]} elements", flush=True)
    assert cells.shape == (1, 160000, 6)
    assert np.abs(cells[0] - sig[None, :]).max() <= 1e-9 * scale
    assert np.abs(nodes[0, :, :6] - sig[None, :]).max() <= 1e-9 * scale
    assert not nodes[0, :, 6:].any()


def test_config3_coupled_sweep_full(ctx):
    """The coupled plate/cavity at full size (1,326,130 DOF): 32 consecutive
    frequencies of the 512-frequency sweep (20-1000 Hz) at the top of the
    range, on 32 slots with the coupled preconditioner; the FP64 residual of
    elPaSo's system is recomputed on the host from elp_dynamic_matrix for
    three of them."""
    m = MG.coupled_plate_cavity()
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(20, 1000, 512)[480:]
    X, rel, it, st = g.sweep(freqs, block=32)
    print(f"[fullsize] coupled sweep relres max {rel.max():.2e} iters {it.min()}-{it.max()} "
          f"refactors {st.refactors}", flush=True)
    assert (rel <= 1e-8).all() and st.unconverged == 0
    rp, ci = g.pattern()
    for k in (0, 17, 31):
        a, F = g.dynamic_matrix(2 * np.pi * freqs[k])
        A = sp.csr_matrix((a, ci, rp), shape=(g.n, g.n))
        res = F - A @ X[k]
        r = np.linalg.norm(res) / np.linalg.norm(F)
        assert r <= 1.5e-8, (freqs[k], r, rel[k])
        # the residual of the diagonally scaled system |D|^-1/2 A |D|^-1/2:
        # the fluid rows (elPaSo's 1/(rho w^2), 1e-6) weigh as much as the
        # plate's (1e9), so the pressure cannot stay behind
        s = 1.0 / np.sqrt(np.abs(A.diagonal()))
        rs = np.linalg.norm(s * res) / np.linalg.norm(s * F)
        print(f"[fullsize] coupled {freqs[k]:.1f} Hz relres {r:.2e} scaled {rs:.2e}", flush=True)
        assert rs <= 1e-7, (freqs[k], rs)
