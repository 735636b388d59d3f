"""The allocation-free node results (elp_node_results_noalloc) against the
allocating entry point (elp_node_results): the same kernels, so the same
bits, on device buffers the caller owns; two calls in a row, the second with
a smaller input, so that the scratch's leftover contents cannot leak into a
result; and a model whose bricks take several scratch-sized chunks of
frequencies."""
import ctypes as C

import numpy as np
import pytest

from elpaso_amd import GpuModel, Material
from elpaso_amd import meshgen as MG

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h = h

    def upload(self, a):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self, p):
        assert self.h.hipFree(p) == 0


def _solutions(n, nf, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nf, n)) + 1j * rng.standard_normal((nf, n))


def _compare(hip, g, freqs, seed):
    """no-alloc nodes == allocating nodes, bit for bit"""
    X = _solutions(g.n, len(freqs), seed)
    _, ref = g.node_results(X, freqs, cells=False)
    nn, _ = g.results_sizes()
    dX = hip.upload(np.ascontiguousarray(X))
    dN = hip.empty(len(freqs) * nn * 9 * 16)
    try:
        g.node_results_noalloc(dX.value, freqs, dN.value)
        got = hip.download(dN, ref)
    finally:
        hip.free(dX)
        hip.free(dN)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert np.abs(ref).max() > 0


def test_noalloc_node_results_equal_allocating(ctx):
    hip = _Hip()
    m = MG.coupled_plate_cavity(n_xy=6, n_plate_z=2, n_cav_z=4, eta=0.02)
    g = GpuModel(ctx, m)
    print("[noalloc] coupled 6x6x(2+4), 5 frequencies", flush=True)
    _compare(hip, g, np.array([20.0, 150.0, 400.0, 730.0, 999.0]), 1)
    print("[noalloc] same model, 2 frequencies", flush=True)
    _compare(hip, g, np.array([55.0, 612.0]), 2)
    h = GpuModel(ctx, MG.acoustic_cavity(4, 3, 3, quadratic=True))
    print("[noalloc] Hex20 cavity 4x3x3, 1 frequency", flush=True)
    _compare(hip, h, np.array([300.0]), 3)


def test_noalloc_node_results_in_chunks(ctx):
    """72,000 bricks: 6.9 MB of centroid stresses per frequency, so 20
    frequencies run as 3 chunks of the 64 MiB scratch."""
    hip = _Hip()
    m = MG.steel_block(60, 60, 20, L=(1.0, 1.0, 0.3))
    m.materials[1] = Material("structure", rho=7850.0, E=210e9, nu=0.3, eta=0.01)
    g = GpuModel(ctx, m)
    print("[noalloc] steel block 60x60x20, 20 frequencies", flush=True)
    _compare(hip, g, np.linspace(10.0, 900.0, 20), 4)
