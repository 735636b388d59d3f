"""GPU harmonic sweep: residual bar and agreement with the direct solve.

Bar: every frequency reaches ||b - A x|| / ||b|| <= 1e-8 in FP64, checked here
independently with the oracle's A(w); solutions agree with SuperLU (the
stand-in for elPaSo's MUMPS LU) to 1e-6 relative (a 1e-8 residual bounds the
error by cond(A) * 1e-8).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle as O
from elpaso_amd import GpuModel
from elpaso_amd import _lib
from elpaso_amd import meshgen as MG

pytestmark = pytest.mark.gpu


def _check_solution(m, freqs, X, rel, rtol=1e-8, direct=True, err_tol=1e-6):
    num = O.Numbering(m)
    pat = O.pattern(m, num)
    for k, f in enumerate(freqs):
        rp, ci, A, F = O.dynamic_matrix(m, 2 * np.pi * f, num, pat)
        Am = O.csr(rp, ci, A, num.n)
        r = np.linalg.norm(F - Am @ X[k]) / np.linalg.norm(F)
        assert r <= 1.5 * rtol, (f, r, rel[k])
        if direct:
            x = sla.splu(Am.tocsc()).solve(F)
            assert np.linalg.norm(X[k] - x) <= err_tol * np.linalg.norm(x), f


@pytest.mark.parametrize("prec,pc", [(_lib.PREC_MIXED, _lib.PC_MG), (_lib.PREC_FP64, _lib.PC_MG),
                                     (_lib.PREC_MIXED, _lib.PC_JACOBI)])
def test_config0_sweep(ctx, prec, pc):
    m = MG.acoustic_cavity()
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(10, 500, 100)[::11]
    print(f"[sweep] cfg0 prec={prec} pc={pc} nf={len(freqs)}", flush=True)
    X, rel, it, st = g.sweep(freqs, precision=prec, precond=pc, block=8)
    assert (rel <= 1e-8).all(), rel
    _check_solution(m, freqs, X, rel)


def test_hex20_sweep_small(ctx):
    m = MG.acoustic_cavity(20, 16, 12, quadratic=True)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([50.0, 400.0, 1200.0, 2000.0])
    X, rel, it, st = g.sweep(freqs, block=4, max_iters=4000)
    assert (rel <= 1e-8).all(), (rel, it)
    _check_solution(m, freqs, X, rel)


def test_hex20_rediscretized_hierarchy_small(ctx):
    """The multigrid path without the separable box solve (Robin face on a
    side the box solve does not cover), at resolved frequencies."""
    m = MG.acoustic_cavity(12, 10, 8, quadratic=True)
    m.box["robin_sides"] |= 1 << 3  # declare an x side: disables the box solve
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([60.0, 250.0])
    X, rel, it, st = g.sweep(freqs, block=4, shift=0.5, smooth=2, max_iters=8000)
    assert (rel <= 1e-8).all(), (rel, it)
    _check_solution(m, freqs, X, rel)


def test_mixed_matches_fp64(ctx):
    m = MG.acoustic_cavity(10, 8, 6)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(10, 500, 16)
    Xm, relm, _, _ = g.sweep(freqs, precision=_lib.PREC_MIXED, block=16)
    Xd, reld, _, _ = g.sweep(freqs, precision=_lib.PREC_FP64)
    assert (relm <= 1e-8).all() and (reld <= 1e-8).all()
    err = np.linalg.norm(Xm - Xd, axis=1) / np.linalg.norm(Xd, axis=1)
    assert (err <= 1e-6).all(), err


def test_block_size_invariance(ctx):
    """A frequency's result does not depend on the block it was batched in
    beyond the refinement tolerance."""
    m = MG.acoustic_cavity(8, 6, 5)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(20, 480, 16)
    X16, r16, _, _ = g.sweep(freqs, block=16)
    X4, r4, _, _ = g.sweep(freqs, block=4)
    err = np.linalg.norm(X16 - X4, axis=1) / np.linalg.norm(X4, axis=1)
    assert (err <= 1e-6).all()


def test_repeat_sweep_bit_identical(ctx):
    m = MG.acoustic_cavity(8, 6, 5)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([100.0, 300.0, 450.0, 480.0])
    X1, _, i1, _ = g.sweep(freqs, block=4)
    X2, _, i2, _ = g.sweep(freqs, block=4)
    assert np.array_equal(X1.view(np.uint64), X2.view(np.uint64)) and np.array_equal(i1, i2)


def test_coupled_sweep_jacobi_small(ctx):
    m = MG.coupled_plate_cavity(n_xy=4, n_plate_z=2, n_cav_z=3)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([20.0, 150.0])
    X, rel, it, st = g.sweep(freqs, precond=_lib.PC_JACOBI, block=4, max_iters=20000)
    assert (rel <= 1e-8).all(), (rel, it)
    # the thin clamped plate makes A(w) ill-conditioned (cond ~ 1e5): a 1e-8
    # residual bounds the error only to ~1e-3
    _check_solution(m, freqs, X, rel, err_tol=1e-3)
    e = _check_fields(m, freqs, X, 1e-5)
    print(f"[sweep] coupled (l1-Jacobi) fields vs scaled direct solve: {e:.2e} iters {it.tolist()}", flush=True)


@pytest.mark.parametrize("quadratic", [False, True])
def test_frequency_alone_equals_mixed(ctx, quadratic):
    """Continuous batching: a frequency solved alone in the block gives the
    same bits as the same frequency solved among others (slots are refilled as
    frequencies converge, columns never mix), and more frequencies than slots
    all reach the bar."""
    m = MG.acoustic_cavity(10, 8, 6, quadratic=quadratic)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(15, 1500, 11)
    print(f"[sweep] alone-vs-mixed quadratic={quadratic}", flush=True)
    Xmix, rmix, imix, _ = g.sweep(freqs, block=4)
    assert (rmix <= 1e-8).all(), rmix
    for k in (0, 5, 10):
        Xa, ra, ia, _ = g.sweep(freqs[k:k + 1], block=4)
        assert np.array_equal(Xa[0].view(np.uint64), Xmix[k].view(np.uint64)), k
        assert ra[0] == rmix[k] and ia[0] == imix[k]
    _check_solution(m, freqs[::5], Xmix[::5], rmix[::5])


@pytest.mark.parametrize("quadratic", [False, True])
def test_sweep_32_slots(ctx, quadratic):
    """32 slots (the row-per-wavefront FP32 operator) with more frequencies
    than slots: every frequency reaches the bar and matches SuperLU."""
    m = MG.acoustic_cavity(12, 10, 8, quadratic=quadratic)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(20, 900, 40)  # resolved on this coarse mesh
    print(f"[sweep] 32 slots quadratic={quadratic}", flush=True)
    X, rel, it, st = g.sweep(freqs, block=32)
    assert (rel <= 1e-8).all(), rel
    _check_solution(m, freqs[::7], X[::7], rel[::7])


def _check_forward_error(m, freqs, X):
    """Small models: FP64 residual of elPaSo's system (relative to the lifted
    right-hand side) at the bar, and the distance to SuperLU within the
    forward-error bound cond(A) relres."""
    num = O.Numbering(m)
    pat = O.pattern(m, num)
    for k, f in enumerate(freqs):
        rp, ci, A, F = O.dynamic_matrix(m, 2 * np.pi * f, num, pat)
        Am = O.csr(rp, ci, A, num.n)
        r = O.relres(m, num, rp, ci, A, F, X[k])
        assert r <= 1.5e-8, (f, r)
        x = sla.splu(Am.tocsc()).solve(F)
        bound = 2 * np.linalg.cond(Am.toarray()) * max(r, 1e-16)
        assert np.linalg.norm(X[k] - x) <= bound * np.linalg.norm(x), (f, bound)


def _check_fields(m, freqs, X, tol):
    """Each field of the solution (displacements, pressure) against a direct
    solve of the symmetrically scaled system (S A S) y = S F, x = S y,
    S = |diag A|^-1/2.  A residual relative to the whole right-hand side
    cannot stand for the pressure of a coupled model: its fluid rows carry
    elPaSo's 1/(rho w^2), 1e-6 against the structure rows' 1e9, so a
    solution with the pressure left at 0 passes a 1e-8 residual bar."""
    import scipy.sparse as sp
    num = O.Numbering(m)
    pat = O.pattern(m, num)
    dof = np.full(num.n, -1)
    nodes, dd = np.nonzero(num.rows >= 0)
    dof[num.rows[nodes, dd]] = dd
    worst = 0.0
    for k, f in enumerate(freqs):
        rp, ci, A, F = O.dynamic_matrix(m, 2 * np.pi * f, num, pat)
        Am = O.csr(rp, ci, A, num.n)
        s = 1.0 / np.sqrt(np.abs(Am.diagonal()))
        S = sp.diags(s)
        x = s * sla.splu((S @ Am @ S).tocsc()).solve(s * F)
        for fld in (dof == O.ELEM_DOFS[O.FLUID8][0], dof != O.ELEM_DOFS[O.FLUID8][0]):
            e = np.linalg.norm(X[k][fld] - x[fld]) / np.linalg.norm(x[fld])
            worst = max(worst, e)
            assert e <= tol, (f, e)
    return worst


@pytest.mark.parametrize("quadratic,nz", [(False, 40), (True, 12)])
def test_impedance_tube_sweep_prescribed_pressure(ctx, quadratic, nz):
    """elp_sweep with a nonzero prescribed pressure (p = 1 on z = 0, lifted
    into the right-hand side): the FP64 residual of elPaSo's system
    (MatZeroRows, F = value) at 1e-8, agreement with SuperLU, and the
    impedance-tube closed form within the element's discretization error."""
    Z = 1.21 * 343.0 * (1 + 0.5j)
    m = MG.impedance_tube(nz, quadratic=quadratic, Z=Z)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([80.0, 300.0, 650.0])
    print(f"[sweep] impedance tube quadratic={quadratic} nz={nz}", flush=True)
    X, rel, it, st = g.sweep(freqs, precond=_lib.PC_JACOBI, block=4, max_iters=20000)
    assert (rel <= 1e-8).all() and st.unconverged == 0, (rel, it)
    _check_forward_error(m, freqs, X)
    num = O.Numbering(m)
    ids = np.sort(m.node_ids)
    z = num.xyz_by_rank(m)[:, 2]
    rows = num.row(ids, 16)
    fixed = np.isclose(z, 0.0)
    for k, f in enumerate(freqs):
        assert np.array_equal(X[k][rows[fixed]], np.ones(int(fixed.sum())))  # exactly the prescribed value
        exact = MG.impedance_tube_field(z, f, Z)
        err = np.abs(X[k][rows] - exact).max() / np.abs(exact).max()
        # about 1.5x the discretization error of the oracle's SuperLU solution
        # of the same mesh (1.1e-4 / 7.5e-3 / 7.0e-2 for Fluid8, 2.0e-7 /
        # 2.2e-4 / 1.1e-2 for Fluid20 at 80 / 300 / 650 Hz)
        lim = {False: {80.0: 2e-4, 300.0: 1.2e-2, 650.0: 0.1}, True: {80.0: 1e-6, 300.0: 4e-4, 650.0: 1.6e-2}}
        assert err < lim[quadratic][f], (f, err)


def test_prescribed_complex_displacements_steel(ctx):
    """Structure rows with complex prescribed displacements on one face, a
    clamped opposite face and a loss factor (no Rayleigh damping: elPaSo's
    system, which the oracle restates, has none)."""
    m = MG.steel_block(6, 5, 4, L=(0.3, 0.25, 0.2), jitter=0.1, rayleigh=(0.0, 0.0))
    from elpaso_amd import Material
    m.materials[1] = Material("structure", rho=7850.0, E=210e9, nu=0.3, eta=0.02)
    xyz = m.xyz
    for nid, p in zip(m.node_ids, xyz):
        if np.isclose(p[0], 0.0):
            for d in (0, 1, 2):
                m.dirichlet.append((int(nid), d, 0.0))
        elif np.isclose(p[0], 0.3):
            m.dirichlet.append((int(nid), 2, 1e-4 * (1 + 0.3j) * (1 + p[1])))
            m.dirichlet.append((int(nid), 0, -2e-5j))
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([100.0, 2500.0])
    print("[sweep] steel block with complex prescribed displacements", flush=True)
    X, rel, it, st = g.sweep(freqs, precond=_lib.PC_JACOBI, block=4, max_iters=40000)
    assert (rel <= 1e-8).all(), (rel, it)
    _check_forward_error(m, freqs, X)


def test_unconverged_sweep_reports_enoconv(ctx):
    """A sweep that cannot reach rtol (one Krylov iteration) returns
    ELP_ENOCONV with its best solutions and residuals filled."""
    m = MG.acoustic_cavity(8, 6, 5)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = np.array([120.0, 480.0])
    L = _lib.load_library()
    o = _lib.default_opts(block=4, max_iters=1)
    X = np.zeros((2, g.n), complex)
    rel = np.zeros(2)
    it = np.zeros(2, np.int32)
    st = _lib.SweepStats()
    rc = L.elp_sweep(g.handle, 2, _lib._ptr(freqs), 0.0, 0.0, _lib.C.byref(o), _lib._ptr(X), _lib._ptr(rel),
                     _lib._ptr(it), _lib.C.byref(st))
    assert rc == _lib.ENOCONV
    assert st.unconverged == 2 and (rel > 1e-8).all() and (rel <= 1.0).all()
    assert np.abs(X).max() > 0
    assert b"did not reach" in L.elp_last_error(None)


@pytest.mark.parametrize("block", [4, 32])
def test_coupled_sweep_block_preconditioner(ctx, block):
    """The coupled plate/cavity with its default preconditioner: the exact
    box solve on the cavity's fluid rows and the direct (block cyclic
    reduction) solve of the clamped plate, factored per window of
    frequencies; FP64 residuals of elPaSo's system at the bar and the
    forward-error bound against SuperLU."""
    m = MG.coupled_plate_cavity(n_xy=10, n_plate_z=2, n_cav_z=12)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(20, 1000, 40)[::3]
    print(f"[sweep] coupled block preconditioner, {len(freqs)} freqs on {block} slots", flush=True)
    X, rel, it, st = g.sweep(freqs, block=block)
    assert (rel <= 1e-8).all() and st.unconverged == 0, (rel, it)
    assert st.refactors >= 1
    _check_forward_error(m, freqs, X)
    e = _check_fields(m, freqs, X, 1e-5)
    print(f"[sweep] coupled fields vs scaled direct solve: {e:.2e}", flush=True)
    # the per-frequency FP64 path (the plate factored at each frequency)
    X64, rel64, _, st64 = g.sweep(freqs, precision=_lib.PREC_FP64)
    assert (rel64 <= 1e-8).all() and st64.unconverged == 0, rel64
    e64 = _check_fields(m, freqs, X64, 1e-5)
    print(f"[sweep] coupled fields, per-frequency FP64: {e64:.2e}", flush=True)


def test_operator_variants_same_bits(ctx):
    """The 32-slot FP32 operator's variants (row kernel, rows grouped 2 or 4
    to a wavefront, LDS row tiles of 16, 32 and 64 rows, tiles with rows
    grouped 2 or 4, tiles with rows grouped 4 and their frequencies split
    over two workgroups) sum the same entries in the same order: the sweeps
    are the same bits, and the set-up's timing choice (op_tile_rows = 0) is
    one of them.  The row kernel runs again after the others (its padding
    rebuilt).  Variant -1 also runs the box cycle's post-smoothing one row
    per wavefront; the others use the post-smoothing kernel the set-up chose
    (grouped half-line tiles where faster): the same bits as well."""
    m = MG.acoustic_cavity(16, 12, 10, quadratic=True)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(10, 2000, 48)
    ref = None
    variants = (-1, -4, 32, 2032, 4032, 4064, 6032, 6064)
    for rows in (-1, -2, -4, 16, 32, 64, 2016, 2032, 4016, 4032, 4064, 6016, 6032, 6064, 0, -1):
        X, rel, it, st = g.sweep(freqs, opts=_lib.default_opts(block=32, op_tile_rows=rows))
        print(f"[sweep] operator variant {rows}: used {st.op_tile_rows}", flush=True)
        assert (rel <= 1e-8).all()
        assert st.op_tile_rows == rows if rows != 0 else st.op_tile_rows in variants
        if ref is None:
            ref = (X, rel, it)
        assert np.array_equal(X.view(np.uint64), ref[0].view(np.uint64)), rows
        assert np.array_equal(rel, ref[1]) and np.array_equal(it, ref[2]), rows


@pytest.mark.parametrize("quadratic,block", [(False, 32), (True, 32), (False, 8)])
def test_graph_replay_same_bits(ctx, quadratic, block):
    """FP32 iterations replayed from the captured HIP graph (with the
    refinement steps and slot refills run eagerly between the replays) give
    the same bits as the iterations launched kernel by kernel."""
    m = MG.acoustic_cavity(12, 10, 8, quadratic=quadratic)
    g = GpuModel(ctx, m)
    g.assemble()
    freqs = MG.linear_frequencies(10, 1500, 40)
    X1, r1, i1, s1 = g.sweep(freqs, opts=_lib.default_opts(block=block, graph=1))
    X2, r2, i2, s2 = g.sweep(freqs, opts=_lib.default_opts(block=block, graph=-1))
    print(f"[sweep] graph replays {s1.graph_replays} of {s1.spmv_calls} iterations", flush=True)
    assert s1.graph_replays > 0 and s2.graph_replays == 0
    assert (r1 <= 1e-8).all()
    assert np.array_equal(X1.view(np.uint64), X2.view(np.uint64))
    assert np.array_equal(r1, r2) and np.array_equal(i1, i2)


def test_coupled_sweep_standin_box(ctx):
    """A coupled plate/cavity whose cavity mesh is no uniform box (jittered
    interior nodes): the coupled preconditioner takes the two-level cycle on
    a stand-in box for the fluid rows (the structure rows still by block
    cyclic reduction); FP64 residuals at the bar and each field against the
    direct solve of the diagonally scaled system."""
    m = MG.coupled_plate_cavity(n_xy=10, n_plate_z=2, n_cav_z=12, jitter=0.2)
    g = GpuModel(ctx, m)
    g.assemble()
    assert g.get_box() is None
    freqs = MG.linear_frequencies(20, 1000, 40)[::3]
    X, rel, it, st = g.sweep(freqs, block=32)
    print(f"[sweep] coupled, stand-in box: iterations {it.tolist()} refactors {st.refactors}", flush=True)
    assert (rel <= 1e-8).all() and st.unconverged == 0, (rel, it)
    assert st.refactors >= 1 and it.mean() < 80
    e = _check_fields(m, freqs, X, 1e-5)
    print(f"[sweep] coupled (stand-in box) fields vs scaled direct solve: {e:.2e}", flush=True)



def test_exact_box_sweeps_unshifted_to_2khz(ctx):
    """Coarse meshes swept to 2 kHz (k h up to 1.5): where the box solve
    inverts the model's own fluid operator (a Hex8 cavity on its box, the
    cavity of a coupled model) the automatic shift is 0, so the iterations
    stay low at the top of the range.  With the k h rule applied there as on
    Hex20 models the same sweeps took 239 (coupled, up to 1,522 at 1.95 kHz)
    and 61 iterations per frequency; the coupled sweep also takes a window's
    frequencies only under that window's structure factorization."""
    freqs = np.sort(np.random.default_rng(3).uniform(10.0, 2000.0, 96))
    g = GpuModel(ctx, MG.acoustic_cavity(26, 21, 16))
    g.assemble()
    X, rel, it, st = g.sweep(freqs, Z=1.21 * 343.0 * (0.5 + 1.0j), block=32, want_x=False)
    print(f"[sweep] Hex8 cavity to 2 kHz: iterations mean {it.mean():.1f} max {it.max()}", flush=True)
    assert (rel <= 1e-8).all() and st.unconverged == 0
    assert it.mean() < 12 and it.max() <= 25, it
    g = GpuModel(ctx, MG.coupled_plate_cavity(n_xy=24, n_plate_z=2, n_cav_z=29))
    g.assemble()
    X, rel, it, st = g.sweep(freqs, block=32, want_x=False)
    print(f"[sweep] coupled 24x24 to 2 kHz: iterations mean {it.mean():.1f} max {it.max()} "
          f"refactors {st.refactors}", flush=True)
    assert (rel <= 1e-8).all() and st.unconverged == 0
    assert it.mean() < 25 and it.max() <= 50, it


def test_long_box_axis_sweep(ctx):
    """A Hex8 duct 140 elements long (2.8 x 0.06 x 0.04 m): the box solve's
    transform along x has 141 points, 9 row blocks of the matrix-core strip
    kernel in FP32 (the Hex20 cavity of the benchmark uses 7); mixed sweep
    on 32 slots against elPaSo's system and SuperLU."""
    m = MG.acoustic_cavity(140, 3, 2, L=(2.8, 0.06, 0.04))
    g = GpuModel(ctx, m)
    g.assemble()
    assert g.get_box() is not None
    freqs = MG.linear_frequencies(10, 2000, 24)
    X, rel, it, st = g.sweep(freqs, block=32)
    print(f"[sweep] duct 140x3x2: iterations mean {it.mean():.1f} max {it.max()}", flush=True)
    assert (rel <= 1e-8).all() and st.unconverged == 0, (rel, it)
    assert it.mean() < 20, it
    _check_forward_error(m, freqs[::6], X[::6])
