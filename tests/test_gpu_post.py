"""GPU element and node results (elp_node_results) against the oracle's
restatement of elPaSo's computeStressesCells / computeStressesNodes:
brick centroid stresses and node sums of the pressure gradient and of the
averaged stresses, on jittered meshes with random solutions.
Bar: max |GPU - oracle| <= 1e-12 max |oracle| per output array (FP64)."""
import numpy as np
import pytest

import oracle as O
from elpaso_amd import BRICK8, BRICK20, FLUID8, FLUID20, ElementGroup, GpuModel, Material, Model
from elpaso_amd import meshgen as MG

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _model(kind, nel, mat, jitter=0.15):
    quad = kind in (FLUID20, BRICK20)
    ids, xyz, eid, conn, _ = MG.box(*nel, L=(1.0, 0.8, 0.6), quadratic=quad, jitter=0.0 if quad else jitter, seed=5)
    m = Model(node_ids=ids, xyz=xyz, name=f"post_{kind}_{nel}")
    m.materials[1] = mat
    m.groups.append(ElementGroup(kind, 1, eid, conn))
    return m


AIR = Material("fluid", rho=1.21, c=343.0)
STEEL_VISCO = Material("structure", rho=7850.0, E=210e9, nu=0.3, eta=0.02, visco_type=2)


def _close(name, got, ref):
    scale = np.abs(ref).max()
    assert scale > 0, name
    err = np.abs(got - ref).max()
    assert err <= TOL * scale, f"{name}: max error {err / scale:.2e} of max |ref|"


def _check(ctx, m, freqs, seed=0):
    print(f"[post] {m.name}", flush=True)
    g = GpuModel(ctx, m)
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((len(freqs), g.n)) + 1j * rng.standard_normal((len(freqs), g.n))
    cells, nodes = g.node_results(X, freqs)
    rc, rn = O.node_results(m, X, freqs)
    assert cells.shape == rc.shape and nodes.shape == rn.shape
    if rc.size:
        _close("cells", cells, rc)
    _close("nodes", nodes, rn)
    return g, X, cells, nodes


@pytest.mark.parametrize("kind,nel,mat", [(FLUID8, (3, 2, 2), AIR), (FLUID20, (2, 2, 1), AIR),
                                          (BRICK8, (3, 2, 2), STEEL_VISCO), (BRICK20, (2, 1, 1), STEEL_VISCO)])
def test_single_kind(ctx, kind, nel, mat):
    _check(ctx, _model(kind, nel, mat), [37.0, 412.5])


def test_coupled_plate_cavity(ctx):
    m = MG.coupled_plate_cavity(n_xy=4, n_plate_z=2, n_cav_z=3)
    _check(ctx, m, [20.0, 250.0, 990.0], seed=1)


def test_linear_pressure_closed_form(ctx):
    m = _model(FLUID8, (4, 3, 2), AIR)
    g = GpuModel(ctx, m)
    num = O.Numbering(m)
    xyz = num.xyz_by_rank(m)
    a = np.array([1.5, -2.0, 0.25])
    X = np.zeros((1, g.n), complex)
    X[0, num.rows[:, 16]] = (xyz @ a + 0.7) * (1 - 0.3j)
    _, nodes = g.node_results(X, [100.0], cells=False)
    cnt = np.zeros(len(xyz), np.int64)
    np.add.at(cnt, num.rank(m.groups[0].conn.ravel()), 1)
    ref = cnt[:, None] * (a * (1 - 0.3j))[None, :]
    assert np.abs(nodes[0, :, 6:] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert not nodes[0, :, :6].any()


def test_many_frequencies_and_repeat_bit_identical(ctx):
    m = _model(BRICK8, (3, 3, 1), STEEL_VISCO)
    g, X, cells, nodes = _check(ctx, m, list(np.linspace(10.0, 1000.0, 40)), seed=2)
    c2, n2 = g.node_results(X, np.linspace(10.0, 1000.0, 40))
    assert np.array_equal(cells, c2) and np.array_equal(nodes, n2)
    c3, n3 = g.node_results(X, np.linspace(10.0, 1000.0, 40), nodes=False)  # cells only
    assert n3 is None and np.array_equal(cells, c3)
