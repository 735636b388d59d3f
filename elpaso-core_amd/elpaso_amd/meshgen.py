"""Seeded synthetic structured meshes in elPaSo conventions.

Node ids are 1-based, x fastest, then y, then z, so ascending id (elPaSo's
node map order) equals the lattice order.  Hex8 connectivity follows the
MSC-Patran order of shape/shapefunctions.cpp:957-1096; Hex20 connectivity
follows the order of shapefunctions.cpp:640-955 (corners 0-3 on zeta=+1,
4-7 on zeta=-1, then 8-19 the midside nodes).
"""
from __future__ import annotations

import math

import numpy as np

from .model import (BRICK8, BRICK20, DOF_X1, DOF_X2, DOF_X3, FLUID8, FLUID20,
                    ElementGroup, Material, Model)

# natural coordinates of the local nodes
HEX8_NAT = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.int64)
HEX20_NAT = np.array([[-1, -1, 1], [-1, 1, 1], [1, 1, 1], [1, -1, 1],
                      [-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1],
                      [-1, 0, 1], [0, 1, 1], [1, 0, 1], [0, -1, 1],
                      [-1, -1, 0], [-1, 1, 0], [1, 1, 0], [1, -1, 0],
                      [-1, 0, -1], [0, 1, -1], [1, 0, -1], [0, -1, -1]], dtype=np.int64)
# Hex8 faces (shape/shapehex8.cpp:119-146), 0-based; normals -y,+y,-z,+x,+z,-x
HEX8_FACES = np.array([[0, 1, 5, 4], [2, 3, 7, 6], [0, 3, 2, 1],
                       [1, 2, 6, 5], [4, 5, 6, 7], [0, 4, 7, 3]], dtype=np.int64)
FACE_PLUS_Z = 4  # 0-based face id with outward normal +z


def _box_lattice(nx, ny, nz, L, origin, order, jitter, seed):
    """Lattice of (order*n+1) points per direction; returns integer lattice
    coordinates of the kept points, their ids and the id lookup array."""
    px, py, pz = order * nx + 1, order * ny + 1, order * nz + 1
    I, J, K = np.meshgrid(np.arange(px), np.arange(py), np.arange(pz), indexing="ij")
    I, J, K = I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()
    if order == 2:  # serendipity: drop face and body centres
        odd = (I % 2) + (J % 2) + (K % 2)
        keep = odd <= 1
    else:
        keep = np.ones(I.shape, dtype=bool)
    I, J, K = I[keep], J[keep], K[keep]
    lookup = -np.ones(px * py * pz, dtype=np.int64)
    lin = I + px * (J + py * K)
    lookup[lin] = np.arange(len(I))
    h = np.array([L[0] / nx, L[1] / ny, L[2] / nz])
    xyz = np.stack([I, J, K], axis=1).astype(np.float64) / order * h + np.asarray(origin, dtype=np.float64)
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        interior = ((I > 0) & (I < px - 1) & (J > 0) & (J < py - 1) & (K > 0) & (K < pz - 1))
        d = rng.uniform(-jitter, jitter, size=(int(interior.sum()), 3)) * h
        xyz[interior] += d
    return I, J, K, xyz, lookup, (px, py, pz)


def box(nx, ny, nz, L=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), quadratic=False,
        first_node_id=1, first_elem_id=1, jitter=0.0, seed=0):
    """Structured hexahedral box mesh.  Returns (node_ids, xyz, elem_ids, conn
    with node ids, element lattice indices (E,3)).  ``jitter`` moves interior
    nodes by U(-jitter, +jitter) times the element edge per coordinate."""
    order = 2 if quadratic else 1
    I, J, K, xyz, lookup, (px, py, pz) = _box_lattice(nx, ny, nz, L, origin, order, jitter, seed)
    node_ids = np.arange(len(I), dtype=np.int64) + first_node_id
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ei, ej, ek = (a.transpose(2, 1, 0).ravel() for a in (ei, ej, ek))
    nat = HEX20_NAT if quadratic else HEX8_NAT
    if quadratic:
        li = (2 * ei + 1)[:, None] + nat[None, :, 0]
        lj = (2 * ej + 1)[:, None] + nat[None, :, 1]
        lk = (2 * ek + 1)[:, None] + nat[None, :, 2]
    else:
        li = ei[:, None] + (nat[None, :, 0] + 1) // 2
        lj = ej[:, None] + (nat[None, :, 1] + 1) // 2
        lk = ek[:, None] + (nat[None, :, 2] + 1) // 2
    conn = lookup[li + px * (lj + py * lk)] + first_node_id
    assert (conn >= first_node_id).all()
    elem_ids = np.arange(len(ei), dtype=np.int64) + first_elem_id
    return node_ids, xyz, elem_ids, conn, np.stack([ei, ej, ek], axis=1)


def nearest_node(node_ids, xyz, p):
    d = np.sum((xyz - np.asarray(p)) ** 2, axis=1)
    return int(node_ids[int(np.argmin(d))])


# ---------------------------------------------------------------------------
# workload builders (sizes are the full workloads by default; tests pass
# smaller element counts)
# ---------------------------------------------------------------------------
AIR = Material("fluid", rho=1.21, c=343.0)
STEEL = Material("structure", rho=7850.0, E=210e9, nu=0.3)


def acoustic_cavity(nx=20, ny=16, nz=12, L=(1.0, 0.8, 0.6), quadratic=False,
                    source=(0.2, 0.3, 0.1), Z=None, lumped=None):
    """Rigid-walled air cavity with a Robin impedance face on z = Lz and a
    unit point source at the node nearest ``source``.  Default Z = rho c (1+0.5i).
    Hex8 faces use the lumped impedance of elPaSo's formula; Hex20 faces the
    consistent face matrix (row-sum lumping of a Quad8 face gives negative
    corner weights)."""
    nid, xyz, eid, conn, lat = box(nx, ny, nz, L, quadratic=quadratic)
    kind = FLUID20 if quadratic else FLUID8
    m = Model(node_ids=nid, xyz=xyz, name=f"cavity_{'hex20' if quadratic else 'hex8'}_{nx}x{ny}x{nz}")
    m.materials[1] = AIR
    m.groups.append(ElementGroup(kind, 1, eid, conn))
    top = lat[:, 2] == nz - 1
    m.imp_elem_ids = eid[top]
    m.imp_faces = np.full(int(top.sum()), FACE_PLUS_Z + 1, dtype=np.int32)
    rho, c = AIR.rho, AIR.c
    m.impedance = complex(rho * c * (1 + 0.5j)) if Z is None else complex(Z)
    m.imp_lumped = (not quadratic) if lumped is None else bool(lumped)
    m.sources.append((nearest_node(nid, xyz, source), 1.0))
    m.box = dict(origin=(0.0, 0.0, 0.0), length=tuple(L), nel=(nx, ny, nz),
                 order=2 if quadratic else 1, robin_sides=1 << FACE_PLUS_Z)
    return m


def irregular_cavity(nx=12, ny=10, nz=8, L=(1.0, 0.8, 0.6), quadratic=False, jitter=0.2, seed=0,
                     l_shape=False, Z=None):
    """An air cavity whose mesh is no uniform box: interior nodes moved by
    U(-jitter, +jitter) element edges, and with l_shape the elements of the
    quarter x > Lx/2, y > Ly/2 removed (an L-shaped floor plan).  Robin face
    (Z, default rho c (1 + 0.5i)) on the top z = Lz, unit source near
    (0.2, 0.3, 0.1); no box description."""
    nid, xyz, eid, conn, lat = box(nx, ny, nz, L, quadratic=quadratic, jitter=jitter, seed=seed)
    keep = np.ones(len(eid), dtype=bool)
    if l_shape:
        keep = ~((lat[:, 0] >= nx // 2) & (lat[:, 1] >= ny // 2))
    eid, conn, lat = eid[keep], conn[keep], lat[keep]
    used = np.unique(conn)
    sel = np.isin(nid, used)
    nid, xyz = nid[sel], xyz[sel]
    kind = FLUID20 if quadratic else FLUID8
    m = Model(node_ids=nid, xyz=xyz,
              name=f"irregular_{'hex20' if quadratic else 'hex8'}_{nx}x{ny}x{nz}{'_L' if l_shape else ''}")
    m.materials[1] = AIR
    m.groups.append(ElementGroup(kind, 1, eid, conn))
    top = lat[:, 2] == nz - 1
    m.imp_elem_ids = eid[top]
    m.imp_faces = np.full(int(top.sum()), FACE_PLUS_Z + 1, dtype=np.int32)
    m.impedance = complex(AIR.rho * AIR.c * (1 + 0.5j)) if Z is None else complex(Z)
    m.imp_lumped = not quadratic
    m.sources.append((nearest_node(nid, xyz, (0.2, 0.3, 0.1)), 1.0))
    return m


def steel_block(nx=200, ny=200, nz=4, L=(1.0, 1.0, 0.02), jitter=0.1, seed=0,
                rayleigh=(1e-5, 2.0)):
    """Linear-elastic steel block of Brick8 elements with seeded jitter of the
    interior nodes (non-affine Jacobians) and Rayleigh damping C = a K + b M.
    The block dimensions are this workload's choice (1.0 x 1.0 x 0.02 m)."""
    nid, xyz, eid, conn, _ = box(nx, ny, nz, L, jitter=jitter, seed=seed)
    m = Model(node_ids=nid, xyz=xyz, name=f"steel_{nx}x{ny}x{nz}")
    m.materials[1] = STEEL
    m.groups.append(ElementGroup(BRICK8, 1, eid, conn))
    m.rayleigh = rayleigh
    return m


def coupled_plate_cavity(n_xy=100, n_plate_z=2, n_cav_z=120, plate_t=0.005,
                         cav_h=1.2, eta=0.01, seed=0, jitter=0.0):
    """Steel plate (Brick8, clamped edges, loss factor eta) lying on top of an
    air cavity (Fluid8).  The plate's bottom face is the cavity's top face
    z = cav_h; the interface Quad4 elements pair coincident fluid and plate
    nodes (separate node ids); orientation +1 means the face normal a1 x a2
    points out of the fluid into the plate.  A unit force F_z = 1 acts at a
    seeded random free node of the plate's top face.  jitter > 0 moves the
    cavity's interior nodes by U(-jitter, +jitter) element edges (its faces
    stay planar): the cavity is then no uniform box and carries no box
    description."""
    fid, fxyz, feid, fconn, _ = box(n_xy, n_xy, n_cav_z, (1.0, 1.0, cav_h), jitter=jitter, seed=seed + 1)
    sid, sxyz, seid, sconn, _ = box(n_xy, n_xy, n_plate_z, (1.0, 1.0, plate_t),
                                    origin=(0.0, 0.0, cav_h),
                                    first_node_id=len(fid) + 1, first_elem_id=len(feid) + 1)
    m = Model(node_ids=np.concatenate([fid, sid]), xyz=np.concatenate([fxyz, sxyz]),
              name=f"coupled_{n_xy}x{n_xy}_{n_plate_z}_{n_cav_z}")
    m.materials[1] = AIR
    m.materials[2] = Material("structure", rho=7850.0, E=210e9, nu=0.3, eta=eta, visco_type=1)
    m.groups.append(ElementGroup(FLUID8, 1, feid, fconn))
    m.groups.append(ElementGroup(BRICK8, 2, seid, sconn))
    # interface quads: fluid top-layer nodes, counter-clockwise seen from +z
    p = n_xy + 1
    top0 = n_cav_z * p * p  # index of the first fluid node on z = cav_h
    i, j = np.meshgrid(np.arange(n_xy), np.arange(n_xy), indexing="ij")
    i, j = i.T.ravel(), j.T.ravel()
    q = np.stack([i + p * j, i + 1 + p * j, i + 1 + p * (j + 1), i + p * (j + 1)], axis=1)
    m.if_fluid_nodes = (top0 + q + 1).astype(np.int64)
    m.if_struct_nodes = (len(fid) + q + 1).astype(np.int64)  # plate bottom layer
    m.if_ids = np.arange(len(q), dtype=np.int64) + 1
    m.if_ori = np.ones(len(q))
    # clamped plate edges
    sx = sxyz[:, 0]
    sy = sxyz[:, 1]
    edge = (np.isclose(sx, 0) | np.isclose(sx, 1.0) | np.isclose(sy, 0) | np.isclose(sy, 1.0))
    for nd in sid[edge]:
        for d in (DOF_X1, DOF_X2, DOF_X3):
            m.dirichlet.append((int(nd), d, 0.0))
    top = np.isclose(sxyz[:, 2], cav_h + plate_t) & ~edge
    rng = np.random.Generator(np.random.PCG64(seed))
    cand = sid[top]
    m.forces.append((int(cand[rng.integers(len(cand))]), 0.0, 0.0, 1.0))
    # the cavity is a uniform Hex8 box: the sweep's coupled preconditioner
    # solves its fluid rows exactly by fast diagonalization
    if jitter == 0:
        m.box = dict(origin=(0.0, 0.0, 0.0), length=(1.0, 1.0, cav_h), nel=(n_xy, n_xy, n_cav_z), order=1,
                     robin_sides=0)
    return m


def impedance_tube(nz, length=1.0, width=0.1, quadratic=False, Z=None, p0=1.0,
                   material=AIR, lumped=None):
    """Air column of 1 x 1 x nz elements along z: the pressure is prescribed
    (Dirichlet, value p0) on the face z = 0 and the face z = length carries a
    Robin impedance Z (default rho c, the matched termination); the side
    walls are rigid.  A plane wave travels along z, so the solution is the
    1-D impedance-tube field p(z) = A e^{-ikz} + B e^{ikz}."""
    nid, xyz, eid, conn, lat = box(1, 1, nz, (width, width, length), quadratic=quadratic)
    kind = FLUID20 if quadratic else FLUID8
    m = Model(node_ids=nid, xyz=xyz, name=f"tube_{'hex20' if quadratic else 'hex8'}_{nz}")
    m.materials[1] = material
    m.groups.append(ElementGroup(kind, 1, eid, conn))
    top = lat[:, 2] == nz - 1
    m.imp_elem_ids = eid[top]
    m.imp_faces = np.full(int(top.sum()), FACE_PLUS_Z + 1, dtype=np.int32)
    rho, c = material.rho, material.c
    m.impedance = complex(rho * c) if Z is None else complex(Z)
    m.imp_lumped = (not quadratic) if lumped is None else bool(lumped)
    for nd in nid[np.isclose(xyz[:, 2], 0.0)]:
        m.dirichlet.append((int(nd), 16, complex(p0)))
    return m


def impedance_tube_field(z, freq, Z, rho=AIR.rho, c=AIR.c, length=1.0, p0=1.0):
    """Closed-form pressure of the impedance tube at heights z (time factor
    e^{+iwt}, the convention of elPaSo's loss factor E(1 + i eta)):
    p = A e^{-ikz} + B e^{ikz}, p(0) = p0, and the Robin condition
    dp/dz = -i w rho p / Z at z = length gives the reflection coefficient
    R = B e^{ikL} / (A e^{-ikL}) = (Z - rho c) / (Z + rho c)."""
    k = 2 * math.pi * freq / c
    R = (Z - rho * c) / (Z + rho * c)
    A = p0 / (1 + R * np.exp(-2j * k * length))
    B = A * R * np.exp(-2j * k * length)
    z = np.asarray(z, dtype=np.float64)
    return A * np.exp(-1j * k * z) + B * np.exp(1j * k * z)


def linear_frequencies(start, stop, steps):
    """elPaSo's list f_k = start + delta k, k < steps (analysisfrequencybasic.cpp:50-55),
    deduplicated and sorted like its std::set."""
    if steps == 1:
        return np.array([float(start)])
    delta = (stop - start) / (steps - 1)
    return np.array(sorted(set(float(start + delta * k) for k in range(steps))))
