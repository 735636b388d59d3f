"""Synthetic workload generators: sizes of the benchmark configurations and
elPaSo conventions (CPU only)."""
import numpy as np

import oracle as O
from elpaso_amd import meshgen as MG


def test_config_sizes():
    m0 = MG.acoustic_cavity()
    assert len(m0.node_ids) == 4641
    # Hex20 100x80x60: corner + edge nodes of the lattice
    nx, ny, nz = 100, 80, 60
    n20 = (nx + 1) * (ny + 1) * (nz + 1) + nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    assert n20 == 1976881
    m = MG.acoustic_cavity(4, 3, 2, quadratic=True)
    assert len(m.node_ids) == 5 * 4 * 3 + 4 * 4 * 3 + 5 * 3 * 3 + 5 * 4 * 2


def test_source_at_requested_point():
    m = MG.acoustic_cavity()
    nid = m.sources[0][0]
    assert np.allclose(m.xyz[nid - 1], (0.2, 0.3, 0.1))
    m2 = MG.acoustic_cavity(10, 8, 6, quadratic=True)
    assert np.allclose(m2.xyz[m2.sources[0][0] - 1], (0.2, 0.3, 0.1))


def test_jitter_interior_only_and_seeded():
    a = MG.steel_block(6, 5, 2, L=(0.03, 0.025, 0.01), jitter=0.1, seed=0)
    b = MG.steel_block(6, 5, 2, L=(0.03, 0.025, 0.01), jitter=0.1, seed=0)
    c = MG.steel_block(6, 5, 2, L=(0.03, 0.025, 0.01), jitter=0.0)
    assert np.array_equal(a.xyz, b.xyz)
    moved = np.any(a.xyz != c.xyz, axis=1)
    on_boundary = np.any(np.isclose(c.xyz, 0) | np.isclose(c.xyz, [0.03, 0.025, 0.01]), axis=1)
    assert moved.any() and not (moved & on_boundary).any()
    assert np.abs(a.xyz - c.xyz).max() <= 0.1 * 0.005 + 1e-15


def test_hex20_connectivity_orientation():
    m = MG.acoustic_cavity(3, 2, 2, quadratic=True)
    conn = m.groups[0].conn
    x = m.xyz[conn[0] - 1]
    nat = MG.HEX20_NAT
    # local node k sits at the natural coordinates of the elPaSo Hex20 order
    assert np.allclose((x - x.min(0)) / (x.max(0) - x.min(0)) * 2 - 1, nat)


def test_coupled_interface_matches_nodes():
    m = MG.coupled_plate_cavity(n_xy=3, n_plate_z=2, n_cav_z=2)
    fx = m.xyz[m.if_fluid_nodes - 1]
    sx = m.xyz[m.if_struct_nodes - 1]
    assert np.allclose(fx, sx)
    n = np.cross(fx[:, 1] - fx[:, 0], fx[:, 3] - fx[:, 0])
    assert np.all(n[:, 2] > 0)  # a1 x a2 points from the fluid into the plate
    num = O.Numbering(m)
    assert num.n == 4 * 4 * 3 + 3 * 4 * 4 * 3


def test_linear_frequencies_match_elpaso_steps():
    f = MG.linear_frequencies(10, 500, 100)
    assert len(f) == 100 and f[0] == 10.0 and abs(f[-1] - 500) < 1e-12
    assert np.allclose(np.diff(f), 490 / 99)
