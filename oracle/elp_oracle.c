/*
 * elp_oracle.c -- CPU restatement of elPaSo-Core's element and assembly
 * arithmetic.  TEST INFRASTRUCTURE ONLY (see elp_oracle.h): the product
 * library never links or calls this file.
 *
 * Paths in citations are relative to elpasoCore/source of elPaSo-Core.
 */
#include "elp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * Gauss-Legendre tables: the 15-significant-digit literals of
 * misc/gausspoints.h:28-45 (the +0.577... point is listed first).
 * ------------------------------------------------------------------------- */
static const double k_gp[10] = {0.0,
                                0.577350269189626,  -0.577350269189626,
                                0.774596669241483,  -0.000000000000000,
                                -0.774596669241483, 0.861136311594053,
                                0.339981043584856,  -0.339981043584856,
                                -0.861136311594053};
static const double k_gw[10] = {2.000000000000000, 1.000000000000000,
                                1.000000000000000, 0.555555555555556,
                                0.888888888888889, 0.555555555555556,
                                0.347854845137454, 0.652145154862546,
                                0.652145154862546, 0.347854845137454};

/* cGaussPoints::Start (gausspoints.cpp:34-48): offset (level-1)*level/2 */
static int start_of(int level) { return (level - 1) * level / 2; }

double orc_gauss_point_1d(int level, int nr) { return k_gp[start_of(level) + nr]; }
double orc_gauss_weight_1d(int level, int nr) { return k_gw[start_of(level) + nr]; }

/* gausspoints.cpp:76-87: r = nr % L^2, z = (nr - r)/L^2, y, x  -> x fastest */
static void split3(int level, int nr, int *x, int *y, int *z) {
  int L2 = level * level;
  int r = nr % L2;
  *z = (nr - r) / L2;
  r = (nr % L2) % level;
  *y = (nr % L2 - r) / level;
  *x = nr - (*y) * level - (*z) * L2;
}

void orc_gauss_point_3d(int level, int nr, double gp[3]) {
  int x, y, z;
  split3(level, nr, &x, &y, &z);
  gp[0] = k_gp[start_of(level) + x];
  gp[1] = k_gp[start_of(level) + y];
  gp[2] = k_gp[start_of(level) + z];
}

/* gausspoints.cpp:184-195: product of the three 1D weights, x*y*z order */
double orc_gauss_weight_3d(int level, int nr) {
  int x, y, z;
  split3(level, nr, &x, &y, &z);
  return k_gw[start_of(level) + x] * k_gw[start_of(level) + y] *
         k_gw[start_of(level) + z];
}

/* gausspoints.cpp:165-174: y = nr % level, x = nr / level (y fastest) */
void orc_gauss_point_2d(int level, int nr, double gp[3]) {
  int y = nr % level, x = (nr - y) / level;
  gp[0] = k_gp[start_of(level) + x];
  gp[1] = k_gp[start_of(level) + y];
  gp[2] = 0.0;
}

double orc_gauss_weight_2d(int level, int nr) {
  int y = nr % level, x = (nr - y) / level;
  return k_gw[start_of(level) + x] * k_gw[start_of(level) + y];
}

/* ---------------------------------------------------------------------------
 * Shape functions.
 * Hex8 node k sits at the natural coordinates of the MSC-Patran order used by
 * shapefunctions.cpp:957-1096: 0(-,-,-) 1(+,-,-) 2(+,+,-) 3(-,+,-) 4(-,-,+)
 * 5(+,-,+) 6(+,+,+) 7(-,+,+).  The trilinear function is evaluated as
 * ((a*b)*c)*0.125 with a = 1 +- xi, which is the reference's operation order.
 * ------------------------------------------------------------------------- */
static const double k_hex8_nat[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1},
                                        {-1, 1, -1},  {-1, -1, 1}, {1, -1, 1},
                                        {1, 1, 1},    {-1, 1, 1}};

/* Hex20 node order of shapefunctions.cpp:640-955 (corners 0-3 on zeta=+1,
 * 4-7 on zeta=-1, then the midside nodes) */
static const double k_hex20_nat[20][3] = {
    {-1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, -1, 1},  {-1, -1, -1},
    {-1, 1, -1}, {1, 1, -1}, {1, -1, -1}, {-1, 0, 1}, {0, 1, 1},
    {1, 0, 1},   {0, -1, 1}, {-1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
    {1, -1, 0},  {-1, 0, -1}, {0, 1, -1}, {1, 0, -1}, {0, -1, -1}};

static double lin(double s, double si) { return si < 0 ? (1. - s) : (1. + s); }

static double hex8(int k, int deriv, const double *g) {
  const double *n = k_hex8_nat[k];
  double a = lin(g[0], n[0]), b = lin(g[1], n[1]), c = lin(g[2], n[2]);
  switch (deriv) {
    case ORC_N: return a * b * c * 0.125;
    case ORC_NXI: return n[0] * (b * c * 0.125);
    case ORC_NETA: return n[1] * (a * c * 0.125);
    default: return n[2] * (a * b * 0.125);
  }
}

/* 20-node serendipity brick: corners 1/8 (1+xi xi_k)(1+eta eta_k)(1+zeta zeta_k)
 * (xi xi_k + eta eta_k + zeta zeta_k - 2); midsides 1/4 (1-s^2)(..)(..). */
static double hex20(int k, int deriv, const double *g) {
  const double *n = k_hex20_nat[k];
  double x = g[0], y = g[1], z = g[2];
  if (n[0] != 0 && n[1] != 0 && n[2] != 0) {
    double a = 1 + x * n[0], b = 1 + y * n[1], c = 1 + z * n[2];
    double s = x * n[0] + y * n[1] + z * n[2] - 2;
    switch (deriv) {
      case ORC_N: return 0.125 * a * b * c * s;
      case ORC_NXI: return 0.125 * n[0] * b * c * (s + a);
      case ORC_NETA: return 0.125 * n[1] * a * c * (s + b);
      default: return 0.125 * n[2] * a * b * (s + c);
    }
  }
  /* midside: the zero coordinate direction is quadratic */
  int q = (n[0] == 0) ? 0 : ((n[1] == 0) ? 1 : 2);
  double f[3], df[3];
  for (int d = 0; d < 3; d++) {
    if (d == q) {
      f[d] = 1 - g[d] * g[d];
      df[d] = -2 * g[d];
    } else {
      f[d] = 1 + g[d] * n[d];
      df[d] = n[d];
    }
  }
  switch (deriv) {
    case ORC_N: return 0.25 * f[0] * f[1] * f[2];
    case ORC_NXI: return 0.25 * df[0] * f[1] * f[2];
    case ORC_NETA: return 0.25 * f[0] * df[1] * f[2];
    default: return 0.25 * f[0] * f[1] * df[2];
  }
}

double orc_shape_hex(int nnod, int k, int deriv, const double gp[3]) {
  return nnod == 8 ? hex8(k, deriv, gp) : hex20(k, deriv, gp);
}

/* Quad4 (shapefunctions.cpp:110-195): 0(-,-) 1(+,-) 2(+,+) 3(-,+);
 * Quad8 serendipity (shapefunctions.cpp:498-): corners then midsides
 * 4(0,-) 5(+,0) 6(0,+) 7(-,0). */
static const double k_quad_nat[8][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1},
                                        {0, -1},  {1, 0},  {0, 1}, {-1, 0}};

double orc_shape_quad(int nnod, int k, int deriv, const double gp[3]) {
  const double *n = k_quad_nat[k];
  double x = gp[0], y = gp[1];
  if (nnod == 4) {
    double a = lin(x, n[0]), b = lin(y, n[1]);
    switch (deriv) {
      case ORC_N: return 0.25 * a * b;
      case ORC_NXI: return n[0] * 0.25 * b;
      default: return n[1] * 0.25 * a;
    }
  }
  if (k < 4) {
    double a = 1 + x * n[0], b = 1 + y * n[1], s = x * n[0] + y * n[1] - 1;
    switch (deriv) {
      case ORC_N: return 0.25 * a * b * s;
      case ORC_NXI: return 0.25 * n[0] * b * (s + a);
      default: return 0.25 * n[1] * a * (s + b);
    }
  }
  if (n[0] == 0) {
    double b = 1 + y * n[1];
    switch (deriv) {
      case ORC_N: return 0.5 * (1 - x * x) * b;
      case ORC_NXI: return -x * b;
      default: return 0.5 * (1 - x * x) * n[1];
    }
  }
  double a = 1 + x * n[0];
  switch (deriv) {
    case ORC_N: return 0.5 * a * (1 - y * y);
    case ORC_NXI: return 0.5 * n[0] * (1 - y * y);
    default: return -y * a;
  }
}

/* cShapeHex8::initializeShapeFunctions (shapehex8.cpp:49-90) */
void orc_hex8_table(double N[256]) {
  for (int n = 0; n < 8; n++) {
    double gp[3];
    orc_gauss_point_3d(2, n, gp);
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 4; d++) N[n * 4 * 8 + d * 8 + k] = hex8(k, d, gp);
  }
}

void orc_hex8_face_table(double Nface[48]) {
  for (int n = 0; n < 4; n++) {
    double gp[3];
    orc_gauss_point_2d(2, n, gp);
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) Nface[n * 3 * 4 + d * 4 + k] = orc_shape_quad(4, k, d, gp);
  }
}

/* shapehex8.cpp:119-146 */
static const int k_hex8_faces[6][4] = {{0, 1, 5, 4}, {2, 3, 7, 6}, {0, 3, 2, 1},
                                       {1, 2, 6, 5}, {4, 5, 6, 7}, {0, 4, 7, 3}};

void orc_hex8_face_nodes(int face, int idx[4]) {
  for (int i = 0; i < 4; i++) idx[i] = k_hex8_faces[face][i];
}

static int hex20_node_at(double x, double y, double z) {
  for (int k = 0; k < 20; k++)
    if (k_hex20_nat[k][0] == x && k_hex20_nat[k][1] == y && k_hex20_nat[k][2] == z) return k;
  return -1;
}

void orc_hex20_face_nodes(int face, int idx[8]) {
  const int *c = k_hex8_faces[face];
  for (int i = 0; i < 4; i++) {
    const double *p = k_hex8_nat[c[i]];
    idx[i] = hex20_node_at(p[0], p[1], p[2]);
  }
  for (int i = 0; i < 4; i++) {
    const double *p = k_hex8_nat[c[i]], *q = k_hex8_nat[c[(i + 1) % 4]];
    idx[4 + i] = hex20_node_at(0.5 * (p[0] + q[0]), 0.5 * (p[1] + q[1]), 0.5 * (p[2] + q[2]));
  }
}

/* ---------------------------------------------------------------------------
 * Jacobian (element/elementfem.cpp:158-210)
 * ------------------------------------------------------------------------- */
static double jac_from_derivs(int nnod, const double *dxi, const double *deta,
                              const double *dzeta, const double *xyz,
                              double J[9], double invJ[9]) {
  for (int i = 0; i < 9; i++) J[i] = 0.0;
  for (int k = 0; k < nnod; k++) {
    const double *x = xyz + 3 * k;
    J[0] += dxi[k] * x[0];
    J[1] += dxi[k] * x[1];
    J[2] += dxi[k] * x[2];
    J[3] += deta[k] * x[0];
    J[4] += deta[k] * x[1];
    J[5] += deta[k] * x[2];
    J[6] += dzeta[k] * x[0];
    J[7] += dzeta[k] * x[1];
    J[8] += dzeta[k] * x[2];
  }
#define A(i, j) J[(i)*3 + (j)]
  double det = A(0, 0) * A(1, 1) * A(2, 2) + A(0, 1) * A(1, 2) * A(2, 0) +
               A(0, 2) * A(1, 0) * A(2, 1) - A(0, 2) * A(1, 1) * A(2, 0) -
               A(0, 0) * A(1, 2) * A(2, 1) - A(0, 1) * A(1, 0) * A(2, 2);
  if (invJ) {
    invJ[0] = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) / det;
    invJ[3] = -(A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) / det;
    invJ[6] = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) / det;
    invJ[1] = -(A(0, 1) * A(2, 2) - A(0, 2) * A(2, 1)) / det;
    invJ[4] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) / det;
    invJ[7] = -(A(0, 0) * A(2, 1) - A(0, 1) * A(2, 0)) / det;
    invJ[2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) / det;
    invJ[5] = -(A(0, 0) * A(1, 2) - A(0, 2) * A(1, 0)) / det;
    invJ[8] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) / det;
  }
#undef A
  return det;
}

/* shape table of one hex element at all Gauss points: T[n][d][k] */
typedef struct {
  int nnod, level, ngp;
  double T[27][4][20];
  double w[27];
} hex_table;

static void make_hex_table(int nnod, hex_table *t) {
  t->nnod = nnod;
  t->level = (nnod == 8) ? 2 : 3;
  t->ngp = t->level * t->level * t->level;
  for (int n = 0; n < t->ngp; n++) {
    double gp[3];
    orc_gauss_point_3d(t->level, n, gp);
    t->w[n] = orc_gauss_weight_3d(t->level, n);
    for (int d = 0; d < 4; d++)
      for (int k = 0; k < nnod; k++) t->T[n][d][k] = orc_shape_hex(nnod, k, d, gp);
  }
}

static const hex_table *table_for(int nnod) {
  static hex_table t8, t20;
  static int init = 0;
  if (!init) {
    make_hex_table(8, &t8);
    make_hex_table(20, &t20);
    init = 1;
  }
  return nnod == 8 ? &t8 : &t20;
}

double orc_jacobian3d(int nnod, int level, int gpnr, const double *xyz,
                      double J[9], double invJ[9]) {
  const hex_table *t = table_for(nnod);
  (void)level;
  return jac_from_derivs(nnod, t->T[gpnr][1], t->T[gpnr][2], t->T[gpnr][3], xyz, J, invJ);
}

/* ---------------------------------------------------------------------------
 * Fluid element (element/fluid/elementfluid3d.cpp)
 * ------------------------------------------------------------------------- */
void orc_fluid_mass(int nnod, const double *xyz, double c, double *MM) {
  const hex_table *t = table_for(nnod);
  const double cf2 = 1.0 / (c * c); /* :58-59 */
  for (int n = 0; n < t->ngp; n++) {
    double J[9];
    double det = jac_from_derivs(nnod, t->T[n][1], t->T[n][2], t->T[n][3], xyz, J, NULL);
    double wdJ = t->w[n] * det;
    for (int z = 0; z < nnod; z++)
      for (int s = 0; s < nnod; s++)
        MM[z * nnod + s] += cf2 * t->T[n][0][z] * t->T[n][0][s] * wdJ; /* :94-96 */
  }
}

void orc_fluid_stiffness(int nnod, const double *xyz, double *KM) {
  const hex_table *t = table_for(nnod);
  double JH[3][20];
  for (int n = 0; n < t->ngp; n++) {
    double J[9], iJ[9];
    double det = jac_from_derivs(nnod, t->T[n][1], t->T[n][2], t->T[n][3], xyz, J, iJ);
    double wdJ = t->w[n] * det;
    /* JH = invJ * H, H rows = dN/dxi, dN/deta, dN/dzeta (:124-137) */
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < nnod; k++) {
        double s = 0.0;
        for (int j = 0; j < 3; j++) s += iJ[i * 3 + j] * t->T[n][1 + j][k];
        JH[i][k] = s;
      }
    /* result = JH^T JH, scaled by wdJ, added (:160-163) */
    for (int z = 0; z < nnod; z++)
      for (int s = 0; s < nnod; s++) {
        double r = 0.0;
        for (int i = 0; i < 3; i++) r += JH[i][z] * JH[i][s];
        KM[z * nnod + s] += r * wdJ;
      }
  }
}

int orc_surface_integral(int nnod_face, const double *xyz, double *C) {
  int level = (nnod_face == 4) ? 2 : 3;
  int ngps = level * level;
  for (int n = 0; n < ngps; n++) {
    double gp[3];
    orc_gauss_point_2d(level, n, gp);
    double w = orc_gauss_weight_2d(level, n);
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0}, N[8];
    for (int k = 0; k < nnod_face; k++) {
      double nx = orc_shape_quad(nnod_face, k, ORC_NXI, gp);
      double ne = orc_shape_quad(nnod_face, k, ORC_NETA, gp);
      N[k] = orc_shape_quad(nnod_face, k, ORC_N, gp);
      for (int d = 0; d < 3; d++) {
        a1[d] += nx * xyz[3 * k + d];
        a2[d] += ne * xyz[3 * k + d];
      }
    }
    double a11 = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
    double a22 = a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2];
    double a12 = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
    double det = sqrt(a11 * a22 - a12 * a12); /* :218 */
    if (det <= 0.0) return -1;
    for (int z = 0; z < nnod_face; z++)
      for (int s = 0; s < nnod_face; s++)
        C[z * nnod_face + s] += N[z] * N[s] * w * det; /* :229-231 */
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * Structural brick (element/structure/linear/brick/elementstructurebrick.cpp)
 * ------------------------------------------------------------------------- */
void orc_setup_c(double complex E, double nu, double complex C[36]) {
  for (int i = 0; i < 36; i++) C[i] = 0;
  const double complex f = E / (1.0 + nu) / (1.0 - 2.0 * nu);
  C[0 * 6 + 0] = f * (1. - nu);
  C[1 * 6 + 1] = f * (1. - nu);
  C[2 * 6 + 2] = f * (1. - nu);
  C[3 * 6 + 3] = f * (1.0 - 2.0 * nu) / 2.0;
  C[4 * 6 + 4] = f * (1.0 - 2.0 * nu) / 2.0;
  C[5 * 6 + 5] = f * (1.0 - 2.0 * nu) / 2.0;
  C[0 * 6 + 1] = f * nu;
  C[1 * 6 + 0] = f * nu;
  C[0 * 6 + 2] = f * nu;
  C[2 * 6 + 0] = f * nu;
  C[1 * 6 + 2] = f * nu;
  C[2 * 6 + 1] = f * nu;
}

double complex orc_e_visco(double E, double eta, int type, double omega) {
  if (type == 2) return E * (1. + I * omega * eta);
  return E * (1. + I * eta);
}

/* infam::BT_C_B_wdJ (math/matrix.h:757-790): BT = B^T * wdJ; res1 = BT*C;
 * res2 = res1*B; KK += res2.  B is nr x ncol real, C nr x nr complex. */
static void bt_c_b_wdj(int nr, int ncol, const double *B, const double complex *C,
                       double wdJ, double complex *KK) {
  double BT[60 * 6];
  double complex res1[60 * 6];
  for (int i = 0; i < ncol; i++)
    for (int j = 0; j < nr; j++) BT[i * nr + j] = B[j * ncol + i] * wdJ;
  for (int i = 0; i < ncol; i++)
    for (int j = 0; j < nr; j++) {
      double complex s = 0;
      for (int k = 0; k < nr; k++) s += BT[i * nr + k] * C[k * nr + j];
      res1[i * nr + j] = s;
    }
  for (int i = 0; i < ncol; i++)
    for (int j = 0; j < ncol; j++) {
      double complex s = 0;
      for (int k = 0; k < nr; k++) s += res1[i * nr + k] * B[k * ncol + j];
      KK[i * ncol + j] += s;
    }
}

void orc_brick_stiffness(int nnod, const double *xyz, double complex E,
                         double nu, double complex *KM) {
  const hex_table *t = table_for(nnod);
  const int ncol = 3 * nnod;
  double complex Cm[36];
  orc_setup_c(E, nu, Cm);
  double *B = calloc(6 * ncol, sizeof(double));
  for (int n = 0; n < t->ngp; n++) {
    double J[9], iJ[9];
    double det = jac_from_derivs(nnod, t->T[n][1], t->T[n][2], t->T[n][3], xyz, J, iJ);
    for (int k = 0; k < nnod; k++) { /* :478-500 */
      double Nx = iJ[0] * t->T[n][1][k] + iJ[1] * t->T[n][2][k] + iJ[2] * t->T[n][3][k];
      double Ny = iJ[3] * t->T[n][1][k] + iJ[4] * t->T[n][2][k] + iJ[5] * t->T[n][3][k];
      double Nz = iJ[6] * t->T[n][1][k] + iJ[7] * t->T[n][2][k] + iJ[8] * t->T[n][3][k];
      B[0 * ncol + 0 + 3 * k] = Nx;
      B[1 * ncol + 1 + 3 * k] = Ny;
      B[2 * ncol + 2 + 3 * k] = Nz;
      B[3 * ncol + 0 + 3 * k] = Ny;
      B[3 * ncol + 1 + 3 * k] = Nx;
      B[4 * ncol + 1 + 3 * k] = Nz;
      B[4 * ncol + 2 + 3 * k] = Ny;
      B[5 * ncol + 0 + 3 * k] = Nz;
      B[5 * ncol + 2 + 3 * k] = Nx;
    }
    double wdJ = t->w[n] * det;
    bt_c_b_wdj(6, ncol, B, Cm, wdJ, KM);
  }
  free(B);
}

void orc_brick_mass(int nnod, const double *xyz, double rho, double *MM) {
  const hex_table *t = table_for(nnod);
  const int ncol = 3 * nnod;
  double complex Cm[9] = {rho, 0, 0, 0, rho, 0, 0, 0, rho};
  double *B = calloc(3 * ncol, sizeof(double));
  double complex *tmp = calloc(ncol * ncol, sizeof(double complex));
  for (int n = 0; n < t->ngp; n++) {
    double J[9];
    double det = jac_from_derivs(nnod, t->T[n][1], t->T[n][2], t->T[n][3], xyz, J, NULL);
    for (int k = 0; k < nnod; k++) {
      B[0 * ncol + 0 + 3 * k] = t->T[n][0][k];
      B[1 * ncol + 1 + 3 * k] = t->T[n][0][k];
      B[2 * ncol + 2 + 3 * k] = t->T[n][0][k];
    }
    bt_c_b_wdj(3, ncol, B, Cm, t->w[n] * det, tmp);
  }
  for (int i = 0; i < ncol * ncol; i++) MM[i] += creal(tmp[i]);
  free(tmp);
  free(B);
}

/* ---------------------------------------------------------------------------
 * Dynamic element matrices
 * ------------------------------------------------------------------------- */
static void face_coords(int nnod, int face, const double *xyz, double *fx, int *idx) {
  int nfn = (nnod == 8) ? 4 : 8;
  if (nnod == 8)
    orc_hex8_face_nodes(face, idx);
  else
    orc_hex20_face_nodes(face, idx);
  for (int k = 0; k < nfn; k++)
    for (int d = 0; d < 3; d++) fx[3 * k + d] = xyz[3 * idx[k] + d];
}

void orc_fluid_dynamic(int nnod, const double *xyz, double rho, double c,
                       double omega, int n_imp, const int *imp_faces,
                       double complex Z, int lumped, double complex *EM) {
  const int nn = nnod * nnod;
  double *MM = calloc(nn, sizeof(double)), *KK = calloc(nn, sizeof(double));
  double complex *KZ = calloc(nn, sizeof(double complex));
  orc_fluid_mass(nnod, xyz, c, MM);
  for (int i = 0; i < nn; i++) MM[i] *= -omega * omega;
  orc_fluid_stiffness(nnod, xyz, KK);
  for (int i = 0; i < nn; i++) KZ[i] = KK[i];
  /* evaluateImpedance, commented formula elementfluid3d.cpp:324-369:
   * factor = w rho i / Z; HD[z] = sum_s C(z,s) factor; KM(z,z) += HD[z] */
  for (int f = 0; f < n_imp; f++) {
    int nfn = (nnod == 8) ? 4 : 8, idx[8];
    double fx[24], C[64];
    memset(C, 0, sizeof C);
    face_coords(nnod, imp_faces[f], xyz, fx, idx);
    orc_surface_integral(nfn, fx, C);
    const double complex factor = omega * rho * I / Z;
    for (int z = 0; z < nfn; z++) {
      if (lumped) {
        double complex HD = 0;
        for (int s = 0; s < nfn; s++) HD += C[z * nfn + s] * factor;
        KZ[idx[z] * nnod + idx[z]] += HD;
      } else {
        for (int s = 0; s < nfn; s++) KZ[idx[z] * nnod + idx[s]] += C[z * nfn + s] * factor;
      }
    }
  }
  /* elementfluid.cpp:71-76: EM += MM; EM += KK; EM *= 1/(rho w^2) */
  const double complex factor = 1. / (rho * omega * omega);
  for (int i = 0; i < nn; i++) EM[i] = (EM[i] + MM[i] + KZ[i]) * factor;
  free(MM);
  free(KK);
  free(KZ);
}

void orc_brick_dynamic(int nnod, const double *xyz, double complex Evis,
                       double nu, double rho, double omega,
                       double complex *EM) {
  const int n = 3 * nnod, nn = n * n;
  double *lM = calloc(nn, sizeof(double));
  double complex *lK = calloc(nn, sizeof(double complex));
  orc_brick_mass(nnod, xyz, rho, lM);
  for (int i = 0; i < nn; i++) lM[i] *= -omega * omega;
  orc_brick_stiffness(nnod, xyz, Evis, nu, lK);
  for (int i = 0; i < nn; i++) EM[i] = EM[i] + lM[i] + lK[i];
  free(lM);
  free(lK);
}

/* ---------------------------------------------------------------------------
 * FSI coupling (element/interface/elementinterfacemindlin.cpp:80-181)
 * ------------------------------------------------------------------------- */
void orc_coupling_matrix(const double *xyz, double *C) {
  const int nnod = 4, ngp = 2;
  for (int n = 0; n < ngp * ngp; n++) {
    double gp[3], N[4], Nxi[4], Neta[4];
    orc_gauss_point_2d(ngp, n, gp);
    double weight = orc_gauss_weight_2d(ngp, n);
    for (int k = 0; k < nnod; k++) {
      N[k] = orc_shape_quad(4, k, ORC_N, gp);
      Nxi[k] = orc_shape_quad(4, k, ORC_NXI, gp);
      Neta[k] = orc_shape_quad(4, k, ORC_NETA, gp);
    }
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    for (int k = 0; k < nnod; k++)
      for (int d = 0; d < 3; d++) {
        a1[d] += Nxi[k] * xyz[3 * k + d];
        a2[d] += Neta[k] * xyz[3 * k + d];
      }
    double a11 = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
    double a22 = a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2];
    double a12 = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
    double det = sqrt(a11 * a22 - a12 * a12);
    for (int z = 0; z < nnod; z++)
      for (int s = 0; s < nnod; s++) C[(3 * z + 2) * 4 + s] += N[z] * N[s] * weight * det;
  }
  /* transformC: nodal normal a1 x a2 at the element node, scaled to 1 */
  static const double xi_n[4] = {-1., +1., +1., -1.}, eta_n[4] = {-1., -1., +1., +1.};
  for (int n = 0; n < nnod; n++) {
    double gp[3] = {xi_n[n], eta_n[n], 0.};
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    for (int k = 0; k < nnod; k++) {
      double nx = orc_shape_quad(4, k, ORC_NXI, gp), ne = orc_shape_quad(4, k, ORC_NETA, gp);
      for (int d = 0; d < 3; d++) {
        a1[d] += nx * xyz[3 * k + d];
        a2[d] += ne * xyz[3 * k + d];
      }
    }
    double nv[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2],
                    a1[0] * a2[1] - a1[1] * a2[0]};
    double len = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    for (int d = 0; d < 3; d++) nv[d] *= 1. / len;
    for (int s = 0; s < nnod; s++) {
      double v = C[(3 * n + 2) * 4 + s];
      C[(3 * n) * 4 + s] = nv[0] * v;
      C[(3 * n + 1) * 4 + s] = nv[1] * v;
      C[(3 * n + 2) * 4 + s] = nv[2] * v;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Global pattern and scatter (analysis/analysis.cpp)
 * ------------------------------------------------------------------------- */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int64_t orc_pattern(int n, int n_elem, int ndpe, const int *edofs, int n_extra,
                    const int *extra_rows, const int *extra_cols, int *rowptr,
                    int *colind) {
  /* getLocalSparsityPattern (:806-867) + preallocateMatrix's diagonal
   * insertion (:790-792): std::set per row == sort + unique */
  int64_t *cnt = calloc((size_t)n + 1, sizeof(int64_t));
  for (int e = 0; e < n_elem; e++)
    for (int a = 0; a < ndpe; a++) cnt[edofs[(int64_t)e * ndpe + a]] += ndpe;
  for (int i = 0; i < n_extra; i++) cnt[extra_rows[i]] += 1;
  for (int r = 0; r < n; r++) cnt[r] += 1; /* diagonal */
  int64_t *off = malloc(((size_t)n + 1) * sizeof(int64_t));
  off[0] = 0;
  for (int r = 0; r < n; r++) off[r + 1] = off[r] + cnt[r];
  int *buf = malloc((size_t)off[n] * sizeof(int));
  int64_t *fill = calloc((size_t)n, sizeof(int64_t));
  for (int e = 0; e < n_elem; e++) {
    const int *d = edofs + (int64_t)e * ndpe;
    for (int a = 0; a < ndpe; a++) {
      int r = d[a];
      for (int b = 0; b < ndpe; b++) buf[off[r] + fill[r]++] = d[b];
    }
  }
  for (int i = 0; i < n_extra; i++) {
    int r = extra_rows[i];
    buf[off[r] + fill[r]++] = extra_cols[i];
  }
  for (int r = 0; r < n; r++) buf[off[r] + fill[r]++] = r;
  int64_t nnz = 0;
  if (rowptr) rowptr[0] = 0;
  for (int r = 0; r < n; r++) {
    int *s = buf + off[r];
    int64_t m = fill[r];
    qsort(s, (size_t)m, sizeof(int), cmp_int);
    int64_t u = 0;
    for (int64_t i = 0; i < m; i++)
      if (i == 0 || s[i] != s[i - 1]) {
        if (colind) colind[nnz + u] = s[i];
        u++;
      }
    nnz += u;
    if (rowptr) rowptr[r + 1] = (int)nnz;
  }
  free(cnt);
  free(off);
  free(buf);
  free(fill);
  return nnz;
}

int64_t orc_find(const int *rowptr, const int *colind, int r, int c) {
  int lo = rowptr[r], hi = rowptr[r + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (colind[mid] == c) return mid;
    if (colind[mid] < c) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

static int nnod_of(int kind) { return (kind == ORC_FLUID8 || kind == ORC_BRICK8) ? 8 : 20; }
static int is_fluid(int kind) { return kind == ORC_FLUID8 || kind == ORC_FLUID20; }

static int elem_coords(int nnod, const int *conn, const double *xyz, double *ex) {
  for (int k = 0; k < nnod; k++)
    for (int d = 0; d < 3; d++) ex[3 * k + d] = xyz[3 * (int64_t)conn[k] + d];
  return 0;
}

static int check_det(int nnod, const double *ex) {
  const hex_table *t = table_for(nnod);
  for (int n = 0; n < t->ngp; n++) {
    double J[9];
    if (jac_from_derivs(nnod, t->T[n][1], t->T[n][2], t->T[n][3], ex, J, NULL) <= 0.0) return -1;
  }
  return 0;
}

/* MatSetValues(ADD_VALUES) of a dense nd x nd element block, row-major
 * (analysis.cpp:1238-1246) */
static void scatter_real(int nd, const int *rows, const double *EM,
                         const int *rowptr, const int *colind, double *A) {
  for (int i = 0; i < nd; i++)
    for (int j = 0; j < nd; j++) {
      int64_t p = orc_find(rowptr, colind, rows[i], rows[j]);
      A[p] += EM[i * nd + j];
    }
}

static void scatter_cplx(int nd, const int *rows, const double complex *EM,
                         const int *rowptr, const int *colind, double complex *A) {
  for (int i = 0; i < nd; i++)
    for (int j = 0; j < nd; j++) {
      int64_t p = orc_find(rowptr, colind, rows[i], rows[j]);
      A[p] += EM[i * nd + j];
    }
}

int orc_assemble_km(int kind, int n_elem, const int *conn, const double *xyz,
                    const int *edofs, double p0, double p1, double p2,
                    int fluid_rho_scale, const int *rowptr, const int *colind,
                    double *K, double *M) {
  const int nnod = nnod_of(kind);
  const int nd = is_fluid(kind) ? nnod : 3 * nnod;
  double ex[60];
  double *Ke = malloc(sizeof(double) * nd * nd), *Me = malloc(sizeof(double) * nd * nd);
  double complex *Kc = malloc(sizeof(double complex) * nd * nd);
  for (int e = 0; e < n_elem; e++) {
    elem_coords(nnod, conn + (int64_t)e * nnod, xyz, ex);
    if (check_det(nnod, ex)) {
      free(Ke); free(Me); free(Kc);
      return -(1 + e);
    }
    memset(Ke, 0, sizeof(double) * nd * nd);
    memset(Me, 0, sizeof(double) * nd * nd);
    if (is_fluid(kind)) {
      orc_fluid_stiffness(nnod, ex, Ke);
      orc_fluid_mass(nnod, ex, p1, Me);
      if (fluid_rho_scale) {
        for (int i = 0; i < nd * nd; i++) {
          Ke[i] /= p0;
          Me[i] /= p0;
        }
      }
    } else {
      memset(Kc, 0, sizeof(double complex) * nd * nd);
      orc_brick_stiffness(nnod, ex, p0, p1, Kc);
      for (int i = 0; i < nd * nd; i++) Ke[i] = creal(Kc[i]);
      orc_brick_mass(nnod, ex, p2, Me);
    }
    const int *rows = edofs + (int64_t)e * nd;
    if (K) scatter_real(nd, rows, Ke, rowptr, colind, K);
    if (M) scatter_real(nd, rows, Me, rowptr, colind, M);
  }
  free(Ke);
  free(Me);
  free(Kc);
  return 0;
}

int orc_assemble_dynamic(int kind, int n_elem, const int *conn,
                         const double *xyz, const int *edofs, double p0,
                         double p1, double p2, double complex Evis,
                         double omega, const unsigned char *imp_mask,
                         double complex Z, int lumped, const int *rowptr,
                         const int *colind, double complex *A) {
  const int nnod = nnod_of(kind);
  const int nd = is_fluid(kind) ? nnod : 3 * nnod;
  double ex[60];
  double complex *EM = malloc(sizeof(double complex) * nd * nd);
  for (int e = 0; e < n_elem; e++) {
    elem_coords(nnod, conn + (int64_t)e * nnod, xyz, ex);
    if (check_det(nnod, ex)) {
      free(EM);
      return -(1 + e);
    }
    memset(EM, 0, sizeof(double complex) * nd * nd);
    if (is_fluid(kind)) {
      int faces[6], nf = 0;
      if (imp_mask)
        for (int f = 0; f < 6; f++)
          if (imp_mask[e] & (1u << f)) faces[nf++] = f;
      orc_fluid_dynamic(nnod, ex, p0, p1, omega, nf, faces, Z, lumped, EM);
    } else {
      orc_brick_dynamic(nnod, ex, Evis, p1, p2, omega, EM);
    }
    scatter_cplx(nd, edofs + (int64_t)e * nd, EM, rowptr, colind, A);
  }
  free(EM);
  return 0;
}

int orc_assemble_face(int nfn, int n_faces, const int *face_nodes,
                      const double *xyz, const int *fdofs, int lumped,
                      const int *rowptr, const int *colind, double *C) {
  double fx[24], Cf[64];
  for (int f = 0; f < n_faces; f++) {
    const int *fn = face_nodes + (int64_t)f * nfn;
    for (int k = 0; k < nfn; k++)
      for (int d = 0; d < 3; d++) fx[3 * k + d] = xyz[3 * (int64_t)fn[k] + d];
    memset(Cf, 0, sizeof Cf);
    if (orc_surface_integral(nfn, fx, Cf)) return -(1 + f);
    const int *rows = fdofs + (int64_t)f * nfn;
    for (int z = 0; z < nfn; z++) {
      if (lumped) {
        double s = 0;
        for (int q = 0; q < nfn; q++) s += Cf[z * nfn + q];
        C[orc_find(rowptr, colind, rows[z], rows[z])] += s;
      } else {
        for (int q = 0; q < nfn; q++) C[orc_find(rowptr, colind, rows[z], rows[q])] += Cf[z * nfn + q];
      }
    }
  }
  return 0;
}

int orc_assemble_fsi(int n_if, const int *fnodes, const double *xyz,
                     const int *srows, const int *frows, const double *ori,
                     const int *rowptr, const int *colind, double *S) {
  double fx[12], C[48];
  for (int e = 0; e < n_if; e++) {
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) fx[3 * k + d] = xyz[3 * (int64_t)fnodes[4 * e + k] + d];
    memset(C, 0, sizeof C);
    orc_coupling_matrix(fx, C);
    const double factor = -1. * ori[e];
    const int *sr = srows + 12 * (int64_t)e, *fr = frows + 4 * (int64_t)e;
    for (int z = 0; z < 12; z++)
      for (int s = 0; s < 4; s++) {
        int64_t p = orc_find(rowptr, colind, sr[z], fr[s]);
        if (p < 0) return -1;
        S[p] += C[z * 4 + s] * factor;
      }
    for (int z = 0; z < 4; z++)
      for (int s = 0; s < 12; s++) {
        int64_t p = orc_find(rowptr, colind, fr[z], sr[s]);
        if (p < 0) return -1;
        S[p] += C[s * 4 + z] * factor;
      }
  }
  return 0;
}
