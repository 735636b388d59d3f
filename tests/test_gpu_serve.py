"""Serving: requests of different sizes submitted together overlap on the GPU
(one stream per slot, admission by HBM budget) and each result is the same,
bit for bit, as the same request solved alone: every request has private
buffers and all kernels are deterministic."""
import numpy as np
import pytest

from elpaso_amd import GpuModel, Server
from elpaso_amd import _lib
from elpaso_amd import meshgen as MG

pytestmark = pytest.mark.gpu


def _requests():
    rng = np.random.default_rng(0)
    reqs = []
    for nel, quad in (((10, 8, 6), False), ((20, 16, 12), False), ((8, 6, 5), True), ((12, 10, 8), False)):
        a, b = rng.uniform(0.2, 2.0, size=2)
        Z = 1.21 * 343.0 * (a + 1j * b)
        nf = int(rng.integers(6, 20))
        freqs = np.sort(rng.uniform(10.0, 1000.0, size=nf))
        reqs.append((nel, quad, Z, freqs))
    return reqs


def _model(ctx, nel, quad, Z):
    m = MG.acoustic_cavity(*nel, quadratic=quad, Z=Z)
    return GpuModel(ctx, m)


@pytest.mark.parametrize("budget_mb,slots", [(4096, 4), (1, 2)])
def test_serve_alone_equals_mixed(ctx, budget_mb, slots):
    reqs = _requests()
    opts = _lib.default_opts(block=8)
    alone = []
    for nel, quad, Z, freqs in reqs:
        g = _model(ctx, nel, quad, Z)
        X, rel, it, _ = g.sweep(freqs, opts=opts)
        assert (rel <= 1e-8).all(), rel
        alone.append((X, rel, it))
        g.close()
    print(f"[serve] budget {budget_mb} MB, {slots} slots, {len(reqs)} requests", flush=True)
    sv = Server(ctx, budget_mb << 20, max_slots=slots)
    jobs = [sv.submit(_model(ctx, nel, quad, Z), freqs, opts=opts) for nel, quad, Z, freqs in reqs]
    for k, j in enumerate(jobs):
        X, rel, it, lat = sv.wait(j)
        Xa, ra, ia = alone[k]
        assert lat > 0
        assert np.array_equal(X.view(np.uint64), Xa.view(np.uint64)), k
        assert np.array_equal(rel, ra) and np.array_equal(it, ia), k
    st = sv.stats()
    assert st["admitted"] == len(reqs) and st["queued"] == 0 and st["bytes_in_use"] == 0
    sv.close()
