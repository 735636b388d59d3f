"""Plain-data description of an elPaSo frequency-response model.

The fields mirror what elPaSo's HDF5 parser builds into ``cMesh``
(misc/parser/femparserinterface.cpp:189-663): nodes with ids and coordinates,
element groups (``/Elements/mtxFemElemGroupN`` with ``ElementType`` and
``MaterialId``), materials, element loads (here: Robin impedance faces),
interface elements (``/InterfaceElements``: id, orientation, fluid nodes,
structure nodes), nodal loads and nodal constraints.

Ids are elPaSo ids (arbitrary positive integers).  elPaSo iterates nodes and
elements in ascending id (``std::map``, fedata/mesh.h:49-81); that order fixes
the DOF numbering and the order in which element matrices are added.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

# element kinds (C-ABI values)
FLUID8, BRICK8, FLUID20, BRICK20 = 1, 2, 3, 4
KIND_NAMES = {"Fluid8": FLUID8, "Brick8": BRICK8, "Fluid20": FLUID20, "Brick20": BRICK20}
NODES_PER_ELEM = {FLUID8: 8, BRICK8: 8, FLUID20: 20, BRICK20: 20}
DOFS_PER_NODE = {FLUID8: 1, BRICK8: 3, FLUID20: 1, BRICK20: 3}

# elPaSo DOF enum values (fedata/basics_dof.h:29-55)
DOF_X1, DOF_X2, DOF_X3, DOF_FLUID = 0, 1, 2, 16


@dataclass
class Material:
    """Fluid: rho, c.  Structure: E, nu, rho; eta > 0 makes it viscoelastic
    (E_vis = E(1+i eta) for visco_type 1, E(1+i w eta) for type 2,
    material/structure/isotrop/linear/viscoelastic/materialstructurevisco.cpp:283-285)."""

    kind: str  # "fluid" | "structure"
    rho: float
    c: float = 0.0
    E: float = 0.0
    nu: float = 0.0
    eta: float = 0.0
    visco_type: int = 1


@dataclass
class ElementGroup:
    kind: int
    material_id: int
    elem_ids: np.ndarray  # int64 [E]
    conn: np.ndarray  # int64 [E, npe] node ids, elPaSo local node order


@dataclass
class Model:
    node_ids: np.ndarray  # int64 [N]
    xyz: np.ndarray  # float64 [N, 3]
    groups: List[ElementGroup] = field(default_factory=list)
    materials: Dict[int, Material] = field(default_factory=dict)
    # Robin impedance faces: element ids and 1-based face numbers (elPaSo
    # element-load convention, elementfluid3d.cpp:259).  One impedance value
    # per model; a sweep may override it.
    imp_elem_ids: Optional[np.ndarray] = None
    imp_faces: Optional[np.ndarray] = None
    impedance: complex = 0.0
    imp_lumped: bool = True
    # FSI interface elements (Quad4): orientation and node ids
    if_ids: Optional[np.ndarray] = None
    if_ori: Optional[np.ndarray] = None
    if_fluid_nodes: Optional[np.ndarray] = None  # int64 [I, 4]
    if_struct_nodes: Optional[np.ndarray] = None  # int64 [I, 4]
    # loads: point sources (node id, Q) and structural point forces
    sources: List[Tuple[int, float]] = field(default_factory=list)
    forces: List[Tuple[int, float, float, float]] = field(default_factory=list)
    # Dirichlet constraints (node id, dof enum, value)
    dirichlet: List[Tuple[int, int, complex]] = field(default_factory=list)
    # Rayleigh damping C = alpha K + beta M (no elPaSo counterpart; closed form)
    rayleigh: Tuple[float, float] = (0.0, 0.0)
    # optional structured-box description (origin, length, nel, order,
    # robin_sides bit mask over the 0-based Hex8 face numbering); enables the
    # multigrid preconditioner of the sweep solver
    box: Optional[dict] = None
    name: str = "model"

    @property
    def n_elements(self) -> int:
        return int(sum(len(g.elem_ids) for g in self.groups))
