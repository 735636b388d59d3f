// Host construction of the quadrature and shape-function tables (see elp_fem.h).
#include "elp_fem.h"

#include <array>
#include <vector>

namespace elp {

const int kHex8Nat[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                            {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
const int kHex20Nat[20][3] = {{-1, -1, 1}, {-1, 1, 1},  {1, 1, 1},   {1, -1, 1},  {-1, -1, -1},
                              {-1, 1, -1}, {1, 1, -1},  {1, -1, -1}, {-1, 0, 1},  {0, 1, 1},
                              {1, 0, 1},   {0, -1, 1},  {-1, -1, 0}, {-1, 1, 0},  {1, 1, 0},
                              {1, -1, 0},  {-1, 0, -1}, {0, 1, -1},  {1, 0, -1},  {0, -1, -1}};
static const int kQuadNat[8][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1},
                                   {0, -1},  {1, 0},  {0, 1}, {-1, 0}};
// Hex8 faces, shape/shapehex8.cpp:119-146 (outward normals -y,+y,-z,+x,+z,-x)
static const int kFace8[6][4] = {{0, 1, 5, 4}, {2, 3, 7, 6}, {0, 3, 2, 1},
                                 {1, 2, 6, 5}, {4, 5, 6, 7}, {0, 4, 7, 3}};

// elPaSo's 15-digit Gauss literals per level
static const double kGP2[2] = {0.577350269189626, -0.577350269189626};
static const double kGW2[2] = {1.000000000000000, 1.000000000000000};
static const double kGP3[3] = {0.774596669241483, -0.000000000000000, -0.774596669241483};
static const double kGW3[3] = {0.555555555555556, 0.888888888888889, 0.555555555555556};

static double lin1(double s, int si) { return si < 0 ? (1. - s) : (1. + s); }

// trilinear brick; products are formed as ((a*b)*c)*0.125 like the reference
static double hex8_fn(int k, int d, const double *g) {
  const int *n = kHex8Nat[k];
  double a = lin1(g[0], n[0]), b = lin1(g[1], n[1]), c = lin1(g[2], n[2]);
  switch (d) {
    case 0: return a * b * c * 0.125;
    case 1: return n[0] * (b * c * 0.125);
    case 2: return n[1] * (a * c * 0.125);
    default: return n[2] * (a * b * 0.125);
  }
}

static double hex20_fn(int k, int d, const double *g) {
  const int *n = kHex20Nat[k];
  if (n[0] && n[1] && n[2]) {
    double a = 1 + g[0] * n[0], b = 1 + g[1] * n[1], c = 1 + g[2] * n[2];
    double s = g[0] * n[0] + g[1] * n[1] + g[2] * n[2] - 2;
    switch (d) {
      case 0: return 0.125 * a * b * c * s;
      case 1: return 0.125 * n[0] * b * c * (s + a);
      case 2: return 0.125 * n[1] * a * c * (s + b);
      default: return 0.125 * n[2] * a * b * (s + c);
    }
  }
  double f[3], df[3];
  for (int i = 0; i < 3; i++) {
    if (n[i] == 0) {
      f[i] = 1 - g[i] * g[i];
      df[i] = -2 * g[i];
    } else {
      f[i] = 1 + g[i] * n[i];
      df[i] = n[i];
    }
  }
  switch (d) {
    case 0: return 0.25 * f[0] * f[1] * f[2];
    case 1: return 0.25 * df[0] * f[1] * f[2];
    case 2: return 0.25 * f[0] * df[1] * f[2];
    default: return 0.25 * f[0] * f[1] * df[2];
  }
}

static double quad_fn(int nn, int k, int d, double x, double y) {
  const int *n = kQuadNat[k];
  if (nn == 4) {
    double a = lin1(x, n[0]), b = lin1(y, n[1]);
    return d == 0 ? 0.25 * a * b : (d == 1 ? n[0] * 0.25 * b : n[1] * 0.25 * a);
  }
  if (k < 4) {
    double a = 1 + x * n[0], b = 1 + y * n[1], s = x * n[0] + y * n[1] - 1;
    return d == 0 ? 0.25 * a * b * s : (d == 1 ? 0.25 * n[0] * b * (s + a) : 0.25 * n[1] * a * (s + b));
  }
  if (n[0] == 0) {
    double b = 1 + y * n[1];
    return d == 0 ? 0.5 * (1 - x * x) * b : (d == 1 ? -x * b : 0.5 * (1 - x * x) * n[1]);
  }
  double a = 1 + x * n[0];
  return d == 0 ? 0.5 * a * (1 - y * y) : (d == 1 ? 0.5 * n[0] * (1 - y * y) : -y * a);
}

static int hex20_at(int x, int y, int z) {
  for (int k = 0; k < 20; k++)
    if (kHex20Nat[k][0] == x && kHex20Nat[k][1] == y && kHex20Nat[k][2] == z) return k;
  return -1;
}

static HexTables build() {
  HexTables t{};
  for (int n = 0; n < 8; n++) {  // x fastest
    int x = n % 2, y = (n / 2) % 2, z = n / 4;
    double g[3] = {kGP2[x], kGP2[y], kGP2[z]};
    t.w8[n] = kGW2[x] * kGW2[y] * kGW2[z];
    for (int d = 0; d < 4; d++)
      for (int k = 0; k < 8; k++) t.h8[n][d][k] = hex8_fn(k, d, g);
  }
  for (int n = 0; n < 27; n++) {
    int x = n % 3, y = (n / 3) % 3, z = n / 9;
    double g[3] = {kGP3[x], kGP3[y], kGP3[z]};
    t.w27[n] = kGW3[x] * kGW3[y] * kGW3[z];
    for (int d = 0; d < 4; d++)
      for (int k = 0; k < 20; k++) t.h20[n][d][k] = hex20_fn(k, d, g);
  }
  for (int n = 0; n < 4; n++) {  // y fastest
    int y = n % 2, x = n / 2;
    t.wq4[n] = kGW2[x] * kGW2[y];
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < 4; k++) t.q4[n][d][k] = quad_fn(4, k, d, kGP2[x], kGP2[y]);
  }
  for (int n = 0; n < 9; n++) {
    int y = n % 3, x = n / 3;
    t.wq9[n] = kGW3[x] * kGW3[y];
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < 8; k++) t.q8[n][d][k] = quad_fn(8, k, d, kGP3[x], kGP3[y]);
  }
  static const double xn[4] = {-1., 1., 1., -1.}, en[4] = {-1., -1., 1., 1.};
  for (int n = 0; n < 4; n++)
    for (int k = 0; k < 4; k++) {
      t.q4n[n][0][k] = quad_fn(4, k, 1, xn[n], en[n]);
      t.q4n[n][1][k] = quad_fn(4, k, 2, xn[n], en[n]);
    }
  for (int f = 0; f < 6; f++) {
    for (int i = 0; i < 4; i++) {
      t.face8[f][i] = kFace8[f][i];
      const int *p = kHex8Nat[kFace8[f][i]];
      t.face20[f][i] = hex20_at(p[0], p[1], p[2]);
    }
    for (int i = 0; i < 4; i++) {
      const int *p = kHex8Nat[kFace8[f][i]], *q = kHex8Nat[kFace8[f][(i + 1) % 4]];
      t.face20[f][4 + i] = hex20_at((p[0] + q[0]) / 2, (p[1] + q[1]) / 2, (p[2] + q[2]) / 2);
    }
  }
  return t;
}

const HexTables &host_tables() {
  static const HexTables t = build();
  return t;
}

std::vector<double> hex_derivs_at(int npe, const std::vector<std::array<double, 3>> &pts) {
  std::vector<double> T(pts.size() * 3 * npe);
  for (size_t a = 0; a < pts.size(); a++)
    for (int d = 0; d < 3; d++)
      for (int f = 0; f < npe; f++)
        T[(a * 3 + d) * npe + f] = npe == 8 ? hex8_fn(f, d + 1, pts[a].data()) : hex20_fn(f, d + 1, pts[a].data());
  return T;
}

std::vector<double> hex_nodal_derivs(int npe) {
  std::vector<std::array<double, 3>> pts(npe);
  for (int a = 0; a < npe; a++)
    for (int i = 0; i < 3; i++) pts[a][i] = npe == 8 ? kHex8Nat[a][i] : kHex20Nat[a][i];
  return hex_derivs_at(npe, pts);
}

}  // namespace elp
