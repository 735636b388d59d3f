// Quadrature and shape-function tables of the hex elements and their faces.
//
// Quadrature: the Gauss-Legendre literals of elPaSo (misc/gausspoints.h:28-45,
// 15 significant digits, +0.577... listed first); 3D points are numbered x
// fastest (misc/gausspoints.cpp:76-87) and 2D points y fastest
// (gausspoints.cpp:165-174).  Shape functions: Hex8 in the MSC-Patran node
// order of shape/shapefunctions.cpp:957-1096, the 20-node serendipity brick in
// the order of shapefunctions.cpp:640-955, Quad4 / Quad8 faces as
// shapefunctions.cpp:110-195 / 498-.  Tables are built once on the host in
// FP64 and copied to device constant memory.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace elp {

// table layout T[gp][deriv][node], deriv 0 = N, 1 = d/dxi, 2 = d/deta, 3 = d/dzeta
struct HexTables {
  double h8[8][4][8];
  double w8[8];
  double h20[27][4][20];
  double w27[27];
  // faces: [gp][deriv 0..2][node]
  double q4[4][3][4];
  double wq4[4];
  double q8[9][3][8];
  double wq9[9];
  // nodal normals of the Quad4 interface element at its 4 nodes need the
  // derivative table at the nodes (elementinterfacemindlin.cpp:131-181)
  double q4n[4][2][4];
  int face8[6][4];
  int face20[6][8];
};

const HexTables &host_tables();

// shape-function derivatives d/dxi, d/deta, d/dzeta of the npe-node brick
// (8 or 20) at the given natural coordinates: T[point][deriv][node]
std::vector<double> hex_derivs_at(int npe, const std::vector<std::array<double, 3>> &pts);
// ... at the element's own nodes (elPaSo evaluates nodal gradients there,
// element/fluid/elementfluid3d.cpp:407-430)
std::vector<double> hex_nodal_derivs(int npe);

// natural coordinates of local nodes
extern const int kHex8Nat[8][3];
extern const int kHex20Nat[20][3];

}  // namespace elp
