"""Serving: requests of different sizes and kinds (acoustic cavities on FP32
slots, a coupled plate/cavity on FP64 slots) submitted together are advanced
in groups whose operator, dot products and vector updates are fused into
single launches; each result is the same, bit for bit, as the same request
solved alone through elp_sweep, because every request keeps its own
partition of each fused launch and all kernels are deterministic."""
import numpy as np
import pytest
import scipy.sparse as sp

from elpaso_amd import GpuModel, Server
from elpaso_amd import _lib
from elpaso_amd import meshgen as MG

pytestmark = pytest.mark.gpu


def _requests():
    rng = np.random.default_rng(0)
    reqs = []
    for kind, nel in (("cavity", (10, 8, 6, False)), ("cavity", (20, 16, 12, False)), ("cavity", (8, 6, 5, True)),
                      ("coupled", (8, 2, 10)), ("cavity", (12, 10, 8, False)), ("coupled", (12, 2, 14))):
        a, b = rng.uniform(0.2, 2.0, size=2)
        Z = 1.21 * 343.0 * (a + 1j * b)
        nf = int(rng.integers(6, 40))
        freqs = np.sort(rng.uniform(10.0, 1000.0, size=nf))
        reqs.append((kind, nel, Z, freqs))
    return reqs


def _model(ctx, kind, nel, Z):
    if kind == "cavity":
        return GpuModel(ctx, MG.acoustic_cavity(*nel[:3], quadratic=nel[3], Z=Z))
    return GpuModel(ctx, MG.coupled_plate_cavity(n_xy=nel[0], n_plate_z=nel[1], n_cav_z=nel[2]))


@pytest.mark.parametrize("budget_mb,slots,block", [(8192, 8, 8), (1, 2, 8), (8192, 8, 32)])
def test_serve_alone_equals_mixed(ctx, budget_mb, slots, block):
    reqs = _requests()
    opts = _lib.default_opts(block=block)
    alone = []
    for kind, nel, Z, freqs in reqs:
        g = _model(ctx, kind, nel, Z)
        X, rel, it, _ = g.sweep(freqs, Z=Z if kind == "cavity" else 0j, opts=opts)
        assert (rel <= 1e-8).all(), rel
        alone.append((X, rel, it))
        g.close()
    print(f"[serve] budget {budget_mb} MB, {slots} resident, {len(reqs)} requests", flush=True)
    sv = Server(ctx, budget_mb << 20, max_slots=slots)
    jobs = [sv.submit(_model(ctx, kind, nel, Z), freqs, Z=Z if kind == "cavity" else 0j, opts=opts)
            for kind, nel, Z, freqs in reqs]
    for k, j in enumerate(jobs):
        X, rel, it, lat = sv.wait(j)
        Xa, ra, ia = alone[k]
        assert lat > 0
        assert np.array_equal(X.view(np.uint64), Xa.view(np.uint64)), k
        assert np.array_equal(rel, ra) and np.array_equal(it, ia), k
    st = sv.stats()
    assert st["admitted"] == len(reqs) and st["queued"] == 0 and st["bytes_in_use"] == 0
    sv.close()


def test_served_request_residual(ctx):
    """A mid-size coupled request served among others: its FP64 residuals
    recomputed on the host from elPaSo's system (elp_dynamic_matrix)."""
    m = MG.coupled_plate_cavity(n_xy=30, n_plate_z=2, n_cav_z=36)
    freqs = MG.linear_frequencies(20, 1000, 64)
    ref = GpuModel(ctx, m)
    ref.assemble()
    rp, ci = ref.pattern()
    sv = Server(ctx, 16 << 30, max_slots=4)
    other = sv.submit(GpuModel(ctx, MG.acoustic_cavity(16, 12, 10)), MG.linear_frequencies(10, 2000, 40), block=32)
    job = sv.submit(GpuModel(ctx, m), freqs, block=32)
    print("[serve] coupled 30x30x(2+36) among others", flush=True)
    X, rel, it, lat = sv.wait(job)
    _, rel2, _, _ = sv.wait(other)
    sv.close()
    assert (rel <= 1e-8).all() and (rel2 <= 1e-8).all()
    for k in (0, 31, 63):
        a, F = ref.dynamic_matrix(2 * np.pi * freqs[k])
        A = sp.csr_matrix((a, ci, rp), shape=(ref.n, ref.n))
        r = np.linalg.norm(F - A @ X[k]) / np.linalg.norm(F)
        assert r <= 1.5e-8, (freqs[k], r)


def test_destroy_with_queued_requests(ctx):
    """More requests than resident slots, server destroyed after the first
    result: the started ones finish, the queued ones are dropped, and
    destroy returns."""
    sv = Server(ctx, 16 << 30, max_slots=1)
    jobs = [sv.submit(GpuModel(ctx, MG.acoustic_cavity(8, 6, 5)), np.array([100.0, 200.0, 300.0]), block=4)
            for _ in range(6)]
    print("[serve] destroy with queued requests", flush=True)
    _, rel, _, _ = sv.wait(jobs[0])
    assert (rel <= 1e-8).all()
    sv.close()
