"""Pins of the oracle's element and node results (oracle.node_results, the
restatement of elPaSo's computeStressesCells / computeStressesNodes) by
closed forms the elements reproduce exactly:
- a linear pressure field on jittered Fluid8 elements (and a quadratic one on
  Fluid20): grad p at every element node is the field's gradient, and a node
  sums it over its elements (no averaging, elementfluid3d.cpp:395-404);
- a linear displacement field on jittered Brick8 elements: constant strain,
  so every centroid stress is D eps, components in the order xx, yy, zz, xy,
  yz, xz (distinct shear strains pin the order), and a node's slots 0..5 are
  the average over its elements; a rigid-body motion gives zero stress;
- the loss factor multiplies the stress by 1 + i eta (type 1) or 1 + i w eta."""
import numpy as np
import pytest

import oracle as O
from elpaso_amd import BRICK8, FLUID20, FLUID8, ElementGroup, Material, Model
from elpaso_amd import meshgen as MG


def _mesh(kind, nel=(2, 2, 1), jitter=0.15, mat=None):
    ids, xyz, eid, conn, _ = MG.box(*nel, L=(1.0, 0.8, 0.6), quadratic=kind in (FLUID20,), jitter=jitter, seed=3)
    m = Model(node_ids=ids, xyz=xyz, name=f"post_{kind}")
    m.materials[1] = mat or (Material("fluid", rho=1.21, c=343.0) if kind in (FLUID8, FLUID20)
                             else Material("structure", rho=7850.0, E=210e9, nu=0.3))
    m.groups.append(ElementGroup(kind, 1, eid, conn))
    return m


def _elements_per_node(m, num):
    cnt = np.zeros(len(num.sorted_ids), np.int64)
    for g in m.groups:
        np.add.at(cnt, num.rank(g.conn.ravel()), 1)
    return cnt


def test_linear_pressure_gradient_fluid8():
    m = _mesh(FLUID8)
    num = O.Numbering(m)
    xyz = num.xyz_by_rank(m)
    a = np.array([1.5, -2.0, 0.25]), np.array([0.5, 3.0, -1.0])
    X = np.zeros((2, num.n), complex)
    rows = num.rows[:, 16]
    for f in range(2):
        X[f, rows] = xyz @ a[f] + 0.7 + 1j * (xyz @ (2 * a[f]) - 0.1)
    cells, nodes = O.node_results(m, X, [100.0, 200.0], num)
    assert cells.shape == (2, 0, 6)
    cnt = _elements_per_node(m, num)
    for f in range(2):
        ref = cnt[:, None] * (a[f] + 2j * a[f])[None, :]
        assert np.abs(nodes[f, :, 6:] - ref).max() <= 1e-12 * np.abs(ref).max()
        assert not nodes[f, :, :6].any()


def test_quadratic_pressure_gradient_fluid20():
    m = _mesh(FLUID20, nel=(2, 1, 1), jitter=0.0)  # straight-sided: the map is trilinear, fields stay quadratic
    num = O.Numbering(m)
    xyz = num.xyz_by_rank(m)
    x, y, z = xyz.T
    X = np.zeros((1, num.n), complex)
    X[0, num.rows[:, 16]] = x * x - 2 * y * z + 3 * z + 0.5j * y * y
    _, nodes = O.node_results(m, X, [50.0], num)
    grad = np.stack([2 * x, -2 * z + 1j * y, -2 * y + 3], axis=1)
    ref = _elements_per_node(m, num)[:, None] * grad
    assert np.abs(nodes[0, :, 6:] - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("case", ["strain", "rigid"])
def test_linear_displacement_brick8(case):
    m = _mesh(BRICK8)
    num = O.Numbering(m)
    xyz = num.xyz_by_rank(m)
    x, y, z = xyz.T
    X = np.zeros((1, num.n), complex)
    if case == "strain":
        # eps_xx = 1e-3, eps_yy = -2e-3, eps_zz = 0.5e-3; shear xy = 4e-4, yz = -6e-4, xz = 9e-4
        u = np.stack([1e-3 * x + 4e-4 * y, -2e-3 * y - 6e-4 * z, 0.5e-3 * z + 9e-4 * x], axis=1)
        eps = np.array([1e-3, -2e-3, 0.5e-3, 4e-4, -6e-4, 9e-4])
    else:
        w = np.array([1e-3, -2e-3, 3e-3])  # rotation + translation
        u = np.cross(w, xyz) + np.array([0.1, 0.2, -0.3])
        eps = np.zeros(6)
    for d in range(3):
        X[0, num.rows[:, d]] = u[:, d] * (1 + 0.5j)
    cells, nodes = O.node_results(m, X, [80.0], num)
    E, nu = 210e9, 0.3
    f = E / ((1 + nu) * (1 - 2 * nu))
    D = np.zeros((6, 6))
    D[:3, :3] = f * nu
    np.fill_diagonal(D[:3, :3], f * (1 - nu))
    D[3, 3] = D[4, 4] = D[5, 5] = f * (1 - 2 * nu) / 2
    sig = D @ eps * (1 + 0.5j)
    scale = np.abs(D @ np.abs(np.array([1e-3, 2e-3, 0.5e-3, 4e-4, 6e-4, 9e-4]))).max()
    assert np.abs(cells[0] - sig[None, :]).max() <= 1e-9 * scale
    assert np.abs(nodes[0, :, :6] - sig[None, :]).max() <= 1e-9 * scale
    assert not nodes[0, :, 6:].any()


@pytest.mark.parametrize("vtype", [1, 2])
def test_loss_factor_scales_stress(vtype):
    eta = 0.01
    mv = _mesh(BRICK8, mat=Material("structure", rho=7850.0, E=210e9, nu=0.3, eta=eta, visco_type=vtype))
    m0 = _mesh(BRICK8)
    num = O.Numbering(m0)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((1, num.n)) + 1j * rng.standard_normal((1, num.n))
    c0, _ = O.node_results(m0, X, [120.0], num)
    cv, _ = O.node_results(mv, X, [120.0], num)
    s = 1 + 1j * (2 * np.pi * 120.0 * eta if vtype == 2 else eta)
    assert np.abs(cv - s * c0).max() <= 1e-12 * np.abs(c0).max()
