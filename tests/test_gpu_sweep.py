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
