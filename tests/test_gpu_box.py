"""The box read from the mesh: a model given without a box description (as
every model read from an elPaSo input file is) gets the box its fluid
elements tile, and with it the same preconditioner, iterations and (to
rounding of the lattice spacing) the same solutions as with the description
given; meshes that do not tile a uniform lattice get none."""
import os
import tempfile

import numpy as np
import pytest

from elpaso_amd import GpuModel, h5
from elpaso_amd import meshgen as MG

from h5_elpaso_writer import write_input

pytestmark = pytest.mark.gpu


def _same_box(got, want):
    assert got is not None and got["detected"]
    assert tuple(got["nel"]) == tuple(want["nel"]) and got["order"] == want["order"]
    assert got["robin_sides"] == want.get("robin_sides", 0)
    assert np.allclose(got["origin"], want["origin"], rtol=0, atol=1e-12)
    assert np.allclose(got["length"], want["length"], rtol=1e-12, atol=0)


def _sweep_pair(ctx, m, freqs, g2=None):
    g1 = GpuModel(ctx, m)
    g1.assemble()
    assert not g1.get_box()["detected"]
    g2 = g2 or GpuModel(ctx, m, box=False)
    g2.assemble()
    _same_box(g2.get_box(), m.box)
    X1, r1, i1, _ = g1.sweep(freqs, block=8)
    X2, r2, i2, _ = g2.sweep(freqs, block=8)
    assert (r1 <= 1e-8).all() and (r2 <= 1e-8).all()
    err = np.linalg.norm(X1 - X2, axis=1) / np.linalg.norm(X1, axis=1)
    print(f"[box] {m.name}: iterations {i1.tolist()} / {i2.tolist()}, max difference {err.max():.1e}", flush=True)
    assert np.abs(i1 - i2).max() <= 1
    assert err.max() <= 1e-7
    g1.close()
    g2.close()


@pytest.mark.parametrize("quadratic", [False, True])
def test_cavity_box_read_from_mesh(ctx, quadratic):
    m = MG.acoustic_cavity(10, 8, 6, quadratic=quadratic)
    _sweep_pair(ctx, m, np.array([50.0, 400.0, 900.0, 1500.0]))


def test_coupled_box_read_from_mesh(ctx):
    m = MG.coupled_plate_cavity(n_xy=10, n_plate_z=2, n_cav_z=12)
    _sweep_pair(ctx, m, MG.linear_frequencies(20, 1000, 40)[::5])


def test_box_read_from_elpaso_input_file(ctx):
    """A rigid-walled cavity (elPaSo-Core's input has no impedance faces)
    written in elPaSo's input layout, which has no box description, and read
    back: the box, the preconditioner and the iterations of the model built
    in memory with its description."""
    m = MG.acoustic_cavity(10, 8, 6)
    m.imp_elem_ids, m.imp_faces, m.impedance = None, None, 0.0
    m.box = dict(m.box, robin_sides=0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cavity.hdf5")
        write_input(path, m)
        g, _, _ = h5.read_model(ctx, path)
    _sweep_pair(ctx, m, np.array([50.0, 370.0, 830.0]), g2=g)


def test_no_box_for_irregular_meshes(ctx):
    # one interior node moved off the lattice
    m = MG.acoustic_cavity(6, 5, 4)
    xyz = np.array(m.xyz, dtype=np.float64)
    c = xyz.mean(axis=0)
    k = int(np.argmin(np.linalg.norm(xyz - c, axis=1)))
    xyz[k, 0] += 0.01
    m.xyz = xyz
    g = GpuModel(ctx, m, box=False)
    assert g.get_box() is None
    g.close()
    # a Robin face inside the fluid is not on a box side: no box either
    m = MG.acoustic_cavity(6, 5, 4)
    m.imp_faces = np.full(len(m.imp_elem_ids), 3, dtype=np.int32)  # -z faces of the top layer
    g = GpuModel(ctx, m, box=False)
    assert g.get_box() is None
    g.close()


def test_box_solve_with_prescribed_pressure(ctx):
    """An impedance tube with the pressure prescribed on its z = 0 face: the
    box read from its mesh gives the box solve, with the prescribed rows left
    out of the correction; the same solutions as l1-Jacobi in a fraction of
    its iterations."""
    from elpaso_amd import _lib
    m = MG.impedance_tube(40, Z=1.21 * 343.0 * (0.7 + 0.2j), p0=1.0)
    g = GpuModel(ctx, m)
    g.assemble()
    assert g.get_box() is not None and g.get_box()["detected"]
    freqs = np.array([80.0, 300.0, 900.0, 1500.0])
    X1, r1, i1, _ = g.sweep(freqs, block=8)
    X2, r2, i2, _ = g.sweep(freqs, opts=_lib.default_opts(block=8, precond=_lib.PC_JACOBI, max_iters=20000))
    print(f"[box] tube iterations: box solve {i1.tolist()}, l1-Jacobi {i2.tolist()}", flush=True)
    assert (r1 <= 1e-8).all() and (r2 <= 1e-8).all()
    assert i1.sum() * 3 < i2.sum()
    err = np.linalg.norm(X1 - X2, axis=1) / np.linalg.norm(X2, axis=1)
    assert err.max() < 1e-4, err
    g.close()


@pytest.mark.parametrize("quadratic,l_shape", [(False, False), (True, False), (False, True)])
def test_standin_box_for_irregular_meshes(ctx, quadratic, l_shape):
    """Fluid meshes that are no box (jittered nodes; an L-shaped floor plan)
    get the two-level cycle on a stand-in box over their bounding box; the
    same solutions as l1-Jacobi, in far fewer iterations."""
    from elpaso_amd import _lib
    m = MG.irregular_cavity(12, 10, 8, quadratic=quadratic, l_shape=l_shape, jitter=0.08 if quadratic else 0.2)
    g = GpuModel(ctx, m)
    g.assemble()
    assert g.get_box() is None
    freqs = np.array([60.0, 400.0, 900.0])
    X1, r1, i1, _ = g.sweep(freqs, block=8)
    X2, r2, i2, _ = g.sweep(freqs, opts=_lib.default_opts(block=8, precond=_lib.PC_JACOBI, max_iters=20000))
    print(f"[box] {m.name}: stand-in box cycle {i1.tolist()} iterations, l1-Jacobi {i2.tolist()}", flush=True)
    assert (r1 <= 1e-8).all() and (r2 <= 1e-8).all()
    assert i1.sum() * 3 < i2.sum()
    err = np.linalg.norm(X1 - X2, axis=1) / np.linalg.norm(X2, axis=1)
    assert err.max() < 1e-5, err
    g.close()
