// oracle.cpp -- CPU FP64 restatement of ICARAT's linear-elastic SIMP
// compliance loop on a voxel hexa8 mesh.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be loaded by
// tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
// and by nothing else.  The product library (icarat2_amd) never links, loads
// or calls it.
//
// Every function cites the reference file:line whose behaviour it restates
// (paths relative to the ICARAT2 source tree).  Nothing here is copied from
// the reference; the algorithms are written out from their mathematics and
// from the documented behaviour of the cited code.
//
// Pinning (what checks each function, see tests/test_oracle_*.py):
//   mesh / DOF numbering ........ closed forms (node id, x-fastest), brute force
//   loads ....................... closed forms (point loads, unit-face val/4)
//   KE0 ......................... exact rational H8 stiffness (tensor 1-D
//                                 integrals), symmetry, 6 rigid modes, volume=1
//   CSR K·u ..................... element-by-element K·u (second formulation)
//   Scaling_CG .................. dense numpy solve on small meshes, textbook PCG
//   compliance .................. c = u^T K u identity
//   sensitivity ................. forward finite differences (reference fdm.cpp)
//   density filter .............. reference KAT test_filter.cpp:63-109, O(N^2)
//                                 brute-force neighbour search
//   threshold ................... reference KAT test_filter.cpp:111-123
//   OC .......................... compiled reference OCSolver.hpp (oracle/_ref)
//   MMA ......................... reference KAT test_opt_solver.cpp:17-73 and the
//                                 compiled reference MMASolver.cpp (oracle/_ref)
//   Gaussian density field ...... project-defined input generator (no reference
//                                 counterpart; checked against the product's
//                                 generator bit for bit)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using i64 = long long;

struct Problem {
  int nx, ny, nz;
  double lx, ly, lz;
  double young1, young2, poisson, penalty;
  i64 numnp, nelx, neq;
  int numeq = -1;
  std::vector<double> X;        // numnp*3 node coordinates
  std::vector<int> conn;        // nelx*8 connectivity
  std::vector<char> hasDir;     // numnp
  std::vector<int> dflag;       // numnp*6
  std::vector<double> dval;     // numnp*6
  std::vector<int> dof;         // numnp*3
  std::vector<double> fext;     // numeq
  double KE0[576];              // element stiffness at young = 1
  double vol_e = 1.0;           // element volume (sum of Gauss jac)
  // assembled K on free DOFs (CSR, rows in dof order, columns ascending)
  std::vector<i64> rowptr;
  std::vector<int> col;
  std::vector<double> val;
  std::vector<double> dens;     // physical densities used for K
  std::vector<double> young;    // SIMP young per element
  bool patternBuilt = false;
};

// ---------------------------------------------------------------- element --
// Trilinear hexa8 shape-function derivatives, local node order
// (-,-,-)(+,-,-)(+,+,-)(-,+,-)(-,-,+)(+,-,+)(+,+,+)(-,+,+)
// (reference basis/shape_function.cpp:403-447).
const int kSign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                         {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

template <class R>
void dNdr_h8(const R r[3], R d[3][8]) {
  for (int a = 0; a < 8; ++a) {
    R s0 = kSign[a][0], s1 = kSign[a][1], s2 = kSign[a][2];
    d[0][a] = 0.125 * s0 * (1.0 + s1 * r[1]) * (1.0 + s2 * r[2]);
    d[1][a] = 0.125 * s1 * (1.0 + s2 * r[2]) * (1.0 + s0 * r[0]);
    d[2][a] = 0.125 * s2 * (1.0 + s0 * r[0]) * (1.0 + s1 * r[1]);
  }
}

// Isotropic 3-D elasticity matrix, Voigt [xx,yy,zz,xy,yz,zx] with engineering
// shears (reference problem/elasticity.cpp:72-93).
template <class R>
void De3(R E, R nu, R D[6][6]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) D[i][j] = 0;
  R lam = E * nu / ((1 + nu) * (1 - 2 * nu));
  R mu = E / (2 * (1 + nu));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[i][j] = lam;
  for (int i = 0; i < 3; ++i) D[i][i] = lam + 2 * mu;
  D[3][3] = D[4][4] = D[5][5] = mu;
}

// Strain-displacement matrix B (6x24, DOF order 3a+d, Voigt [xx,yy,zz,xy,yz,zx]
// with engineering shears) at Gauss point ip of the 2x2x2 rule, points
// (+-1/sqrt3)^3 in the listed order, weight 1 (reference basis/bmatrix.hpp:
// 90-162, basis/gauss_integration.cpp:268-289).  Returns jac = det J * weight.
template <class R>
R element_B(const double Xe[8][3], int ip, R B[6][24]) {
  const R g = 1 / std::sqrt((R)3);
  const R pts[8][3] = {{-g, -g, -g}, {g, -g, -g}, {g, g, -g}, {-g, g, -g},
                       {-g, -g, g},  {g, -g, g},  {g, g, g},  {-g, g, g}};
  R dr[3][8];
  dNdr_h8(pts[ip], dr);
  R J[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R s = 0;
      for (int a = 0; a < 8; ++a) s += dr[i][a] * (R)Xe[a][j];
      J[i][j] = s;
    }
  R det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  R Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  R dX[3][8];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 8; ++a)
      dX[i][a] = Ji[i][0] * dr[0][a] + Ji[i][1] * dr[1][a] + Ji[i][2] * dr[2][a];
  for (int i = 0; i < 6; ++i)
    for (int c = 0; c < 24; ++c) B[i][c] = 0;
  for (int a = 0; a < 8; ++a) {
    B[0][3 * a] = dX[0][a];
    B[1][3 * a + 1] = dX[1][a];
    B[2][3 * a + 2] = dX[2][a];
    B[3][3 * a] = dX[1][a];
    B[3][3 * a + 1] = dX[0][a];
    B[4][3 * a + 1] = dX[2][a];
    B[4][3 * a + 2] = dX[1][a];
    B[5][3 * a] = dX[2][a];
    B[5][3 * a + 2] = dX[0][a];
  }
  return det;
}

// Ke = sum_ip B^T De B jac over the 2x2x2 Gauss rule (reference
// problem/elasticity.cpp:211-255).  Returns the element volume (sum of jac).
// The rule integrates the trilinear element exactly, so Ke is evaluated in
// extended precision (long double) and rounded once: every entry is the FP64
// rounding of the exact integral (equal entries of the exact matrix stay
// bit-identical), which makes the operator the same, bit for bit, for any
// implementation that rounds the exact element matrix.
double element_ke(const double Xe[8][3], double E, double nu, double Ke[576]) {
  using R = long double;
  R D[6][6];
  De3<R>(E, nu, D);
  R acc[576];
  for (int i = 0; i < 576; ++i) acc[i] = 0;
  R vol = 0;
  for (int ip = 0; ip < 8; ++ip) {
    R B[6][24];
    const R jac = element_B<R>(Xe, ip, B);
    R DB[6][24];
    for (int i = 0; i < 6; ++i)
      for (int c = 0; c < 24; ++c) {
        R s = 0;
        for (int k = 0; k < 6; ++k) s += D[i][k] * B[k][c];
        DB[i][c] = s;
      }
    for (int r = 0; r < 24; ++r)
      for (int c = 0; c < 24; ++c) {
        R s = 0;
        for (int k = 0; k < 6; ++k) s += B[k][r] * DB[k][c];
        acc[r * 24 + c] += s * jac;
      }
    vol += jac;
  }
  // an entry that vanishes exactly comes out at the rounding level of the
  // extended sums (1e-19 of the largest); every other exact entry is a ratio
  // of small integers times lam, mu and the element sizes, far above 1e-12
  R amax = 0;
  for (int i = 0; i < 576; ++i) amax = std::max(amax, std::fabs(acc[i]));
  for (int i = 0; i < 576; ++i) Ke[i] = std::fabs(acc[i]) < 1e-12L * amax ? 0.0 : (double)acc[i];
  return (double)vol;
}

// x^p of the SIMP law: for p = 3 the correctly rounded cube (the two products
// with their exact error terms), what a correctly rounded std::pow returns;
// otherwise std::pow.
inline double simp_pow(double x, double p) {
  if (p == 3.0) {
    const double s = x * x;
    const double es = std::fma(x, x, -s);
    const double t = s * x;
    const double et = std::fma(s, x, -t);
    return t + std::fma(es, x, et);
  }
  return std::pow(x, p);
}

// SIMP interpolation (reference problem/elasticity.hpp:13-17).
inline double simp(double y1, double y2, double x, double p) {
  const double xp = simp_pow(x, p);
  return (1.0 - xp) * y1 + xp * y2;
}

void element_X(const Problem &P, i64 e, double Xe[8][3]) {
  for (int a = 0; a < 8; ++a)
    for (int d = 0; d < 3; ++d) Xe[a][d] = P.X[3 * (i64)P.conn[8 * e + a] + d];
}

// ----------------------------------------------------------------- assembly --
// Free-DOF K in CSR.  Row (node n, comp i) holds the free DOFs of the 27
// lattice neighbours; ascending column order equals (dz,dy,dx,comp) order
// because free DOFs are numbered in node order
// (reference analysis/linear.hpp:117-166, misc/assembling.hpp:18-39).
void build_pattern(Problem &P) {
  const int NX = P.nx + 1, NY = P.ny + 1, NZ = P.nz + 1;
  P.rowptr.assign((size_t)P.numeq + 1, 0);
  // count
  for (int k = 0; k < NZ; ++k)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        i64 n = i + (i64)NX * (j + (i64)NY * k);
        int cnt = 0;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              int a = i + dx, b = j + dy, c = k + dz;
              if (a < 0 || b < 0 || c < 0 || a >= NX || b >= NY || c >= NZ) continue;
              i64 m = a + (i64)NX * (b + (i64)NY * c);
              for (int d = 0; d < 3; ++d) cnt += P.dof[3 * m + d] < P.numeq;
            }
        for (int d = 0; d < 3; ++d) {
          int r = P.dof[3 * n + d];
          if (r < P.numeq) P.rowptr[r + 1] = cnt;
        }
      }
  for (int r = 0; r < P.numeq; ++r) P.rowptr[r + 1] += P.rowptr[r];
  P.col.resize(P.rowptr[P.numeq]);
  P.val.resize(P.rowptr[P.numeq]);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < NZ; ++k)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        i64 n = i + (i64)NX * (j + (i64)NY * k);
        for (int d = 0; d < 3; ++d) {
          int r = P.dof[3 * n + d];
          if (r >= P.numeq) continue;
          i64 q = P.rowptr[r];
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                int a = i + dx, b = j + dy, c = k + dz;
                if (a < 0 || b < 0 || c < 0 || a >= NX || b >= NY || c >= NZ) continue;
                i64 m = a + (i64)NX * (b + (i64)NY * c);
                for (int dd = 0; dd < 3; ++dd)
                  if (P.dof[3 * m + dd] < P.numeq) P.col[q++] = P.dof[3 * m + dd];
              }
        }
      }
  P.patternBuilt = true;
}

// local node index of lattice offset (ox,oy,oz) in {0,1}^3 inside an element
inline int lnode(int ox, int oy, int oz) {
  static const int t[2][2] = {{0, 3}, {1, 2}};  // [ox][oy]
  return t[ox][oy] + 4 * oz;
}

void assemble_values(Problem &P) {
  const int NX = P.nx + 1, NY = P.ny + 1, NZ = P.nz + 1;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < NZ; ++k)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        i64 n = i + (i64)NX * (j + (i64)NY * k);
        for (int d = 0; d < 3; ++d) {
          int r = P.dof[3 * n + d];
          if (r >= P.numeq) continue;
          i64 q0 = P.rowptr[r];
          // slot table: neighbour (dx,dy,dz,dd) -> position in the row
          int slot[27][3];
          i64 q = q0;
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                int s = (dx + 1) + 3 * ((dy + 1) + 3 * (dz + 1));
                int a = i + dx, b = j + dy, c = k + dz;
                for (int dd = 0; dd < 3; ++dd) slot[s][dd] = -1;
                if (a < 0 || b < 0 || c < 0 || a >= NX || b >= NY || c >= NZ) continue;
                i64 m = a + (i64)NX * (b + (i64)NY * c);
                for (int dd = 0; dd < 3; ++dd)
                  if (P.dof[3 * m + dd] < P.numeq) slot[s][dd] = (int)(q++ - q0);
              }
          for (i64 t = q0; t < P.rowptr[r + 1]; ++t) P.val[t] = 0.0;
          // elements around node n in ascending element id order
          for (int ez = -1; ez <= 0; ++ez)
            for (int ey = -1; ey <= 0; ++ey)
              for (int ex = -1; ex <= 0; ++ex) {
                int a0 = i + ex, b0 = j + ey, c0 = k + ez;
                if (a0 < 0 || b0 < 0 || c0 < 0 || a0 >= P.nx || b0 >= P.ny || c0 >= P.nz)
                  continue;
                i64 e = a0 + (i64)P.nx * (b0 + (i64)P.ny * c0);
                double E = P.young[e];
                int la = lnode(-ex, -ey, -ez);
                for (int bz = 0; bz <= 1; ++bz)
                  for (int by = 0; by <= 1; ++by)
                    for (int bx = 0; bx <= 1; ++bx) {
                      int lb = lnode(bx, by, bz);
                      int s = (ex + bx + 1) + 3 * ((ey + by + 1) + 3 * (ez + bz + 1));
                      for (int dd = 0; dd < 3; ++dd) {
                        if (slot[s][dd] < 0) continue;
                        P.val[q0 + slot[s][dd]] += E * P.KE0[(3 * la + d) * 24 + 3 * lb + dd];
                      }
                    }
              }
        }
      }
}

void spmv(const Problem &P, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < P.numeq; ++r) {
    double s = 0.0;
    for (i64 t = P.rowptr[r]; t < P.rowptr[r + 1]; ++t) s += P.val[t] * x[P.col[t]];
    y[r] = s;
  }
}

double dot(const double *a, const double *b, int n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// --------------------------------------------------------------- filter ----
struct Filter {
  const Problem *P;
  double radius;
  i64 n;
  std::vector<i64> rowptr;   // CSR neighbour lists, ascending element id
  std::vector<int> ids;
  std::vector<double> dis;
  std::vector<double> aa;    // cached normaliser (dfilters_ds)
};

void element_centre(const Problem &P, i64 e, double c[3]) {
  c[0] = c[1] = c[2] = 0.0;
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < 8; ++a) c[d] += P.X[3 * (i64)P.conn[8 * e + a] + d] / 8;
}

// O(N^2) search over all element pairs, exactly as the reference does
// (reference optimization/filter.hpp:254-297).
void neighbours_brute(Filter &F) {
  const Problem &P = *F.P;
  std::vector<std::vector<int>> I(P.nelx);
  std::vector<std::vector<double>> Dv(P.nelx);
  std::vector<double> cen(3 * P.nelx);
  for (i64 e = 0; e < P.nelx; ++e) element_centre(P, e, &cen[3 * e]);
#pragma omp parallel for schedule(dynamic, 16)
  for (i64 i = 0; i < P.nelx; ++i) {
    for (i64 j = 0; j < P.nelx; ++j) {
      double d0 = cen[3 * i] - cen[3 * j], d1 = cen[3 * i + 1] - cen[3 * j + 1],
             d2 = cen[3 * i + 2] - cen[3 * j + 2];
      double dist = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      if (dist <= F.radius) {
        I[i].push_back((int)j);
        Dv[i].push_back(dist);
      }
    }
  }
  F.rowptr.assign(P.nelx + 1, 0);
  for (i64 i = 0; i < P.nelx; ++i) F.rowptr[i + 1] = F.rowptr[i] + (i64)I[i].size();
  F.ids.resize(F.rowptr[P.nelx]);
  F.dis.resize(F.rowptr[P.nelx]);
  for (i64 i = 0; i < P.nelx; ++i) {
    std::copy(I[i].begin(), I[i].end(), F.ids.begin() + F.rowptr[i]);
    std::copy(Dv[i].begin(), Dv[i].end(), F.dis.begin() + F.rowptr[i]);
  }
}

// Structured-lattice neighbour list: the same set and order as the O(N^2)
// search on a uniform voxel grid (ascending element id == (dz,dy,dx) order).
void neighbours_structured(Filter &F) {
  const Problem &P = *F.P;
  const double hx = P.lx / P.nx, hy = P.ly / P.ny, hz = P.lz / P.nz;
  int R = (int)std::ceil(F.radius / std::min(hx, std::min(hy, hz))) + 1;
  F.rowptr.assign(P.nelx + 1, 0);
  std::vector<std::vector<int>> I(P.nelx);
  std::vector<std::vector<double>> Dv(P.nelx);
  std::vector<double> cen(3 * P.nelx);
  for (i64 e = 0; e < P.nelx; ++e) element_centre(P, e, &cen[3 * e]);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < P.nz; ++k)
    for (int j = 0; j < P.ny; ++j)
      for (int i = 0; i < P.nx; ++i) {
        i64 e = i + (i64)P.nx * (j + (i64)P.ny * k);
        for (int dz = -R; dz <= R; ++dz)
          for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
              int a = i + dx, b = j + dy, c = k + dz;
              if (a < 0 || b < 0 || c < 0 || a >= P.nx || b >= P.ny || c >= P.nz) continue;
              i64 f = a + (i64)P.nx * (b + (i64)P.ny * c);
              double d0 = cen[3 * e] - cen[3 * f], d1 = cen[3 * e + 1] - cen[3 * f + 1],
                     d2 = cen[3 * e + 2] - cen[3 * f + 2];
              double dist = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
              if (dist <= F.radius) {
                I[e].push_back((int)f);
                Dv[e].push_back(dist);
              }
            }
      }
  for (i64 i = 0; i < P.nelx; ++i) F.rowptr[i + 1] = F.rowptr[i] + (i64)I[i].size();
  F.ids.resize(F.rowptr[P.nelx]);
  F.dis.resize(F.rowptr[P.nelx]);
  for (i64 i = 0; i < P.nelx; ++i) {
    std::copy(I[i].begin(), I[i].end(), F.ids.begin() + F.rowptr[i]);
    std::copy(Dv[i].begin(), Dv[i].end(), F.dis.begin() + F.rowptr[i]);
  }
}

// Density filter and its chain rule (reference optimization/filter.hpp:106-155).
int density_filter(Filter &F, const double *in, double *out, int isSens) {
  int bad = 0;
  if (!isSens) {
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (i64 i = 0; i < F.n; ++i) {
      double a = 0.0, b = 0.0;
      for (i64 t = F.rowptr[i]; t < F.rowptr[i + 1]; ++t) {
        double w = F.radius - F.dis[t];
        a += w;
        b += w * in[F.ids[t]];
      }
      if (a < 1.0e-10) bad += 1;
      out[i] = b / a;
      F.aa[i] = a;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < F.n; ++i) {
      double s = 0.0;
      for (i64 t = F.rowptr[i]; t < F.rowptr[i + 1]; ++t) {
        double w = F.radius - F.dis[t];
        s += w * in[F.ids[t]] / F.aa[F.ids[t]];
      }
      out[i] = s;
    }
  }
  return bad ? -1 : 0;
}

// Heaviside projection and its derivative (reference optimization/filter.hpp:216-248).
void threshold(i64 n, const double *x, double beta, double T, double *out) {
  for (i64 i = 0; i < n; ++i)
    out[i] = 0.5 * (std::tanh(T * beta) + std::tanh(beta * (x[i] - T))) / std::tanh(T * beta);
}
void threshold_deriv(i64 n, const double *x, double beta, double T, double *out) {
  for (i64 i = 0; i < n; ++i)
    out[i] = (0.5 * (beta * (1.0 - std::pow(std::tanh(beta * (x[i] - T)), 2.0))) /
              std::tanh(T * beta));
}

// ------------------------------------------------------- Helmholtz filter --
// Helmholtz PDE filter (reference optimization/filter_helm.hpp:76-257; B. S.
// Lazarov and O. Sigmund, Filters in topology optimization based on
// Helmholtz-type differential equations, 2010).  On every element
//   Ke = sum_ip (N N^T + c B^T B) jac,  c = (r / (2 sqrt 3))^2   (:191-223)
//   fe = sum_ip x_e N jac, divided by ipmax when sensFlag          (:215-221)
// with B the gradient matrix dN/dX (bmatrix.hpp:52-58, :111-117), assembled
// over all nodes with no boundary condition, K val = F solved by Jacobi-PCG
// from zero to a relative residual tol ("Eigen_CG", analysis/solver/
// solvers.cpp:63-91, default tol 1e-10 solvers.hpp:89), and per element
//   out_e = sum_ip N(ip) . val_e, divided by ipmax unless sensFlag (:229-257).
// dim 2: quad4 with the 2x2 rule in the reference's point order
// (basis/gauss_integration.cpp:72-88), dim 3: hexa8 with 2x2x2 (:268-289).
// Nodes are numbered x fastest, elements x fastest, local nodes
// counter-clockwise on the bottom face then the top (io/mesh_icarat.hpp:480-511).
struct Helm {
  int dim = 3, nx = 1, ny = 1, nz = 1;
  double h[3] = {1, 1, 1};
  double radius = 0.0;
  i64 numnp = 0, nelx = 0;
  int ne = 8, nip = 8;
  std::vector<int> conn;     // nelx * ne
  std::vector<double> ke;    // nelx * ne * ne
  std::vector<double> nj;    // nelx * nip * ne: N_a(ip) jac(ip)
  std::vector<double> nq;    // nip * ne: N_a(ip) (the same on every element)
  std::vector<double> diag;  // numnp
  int last_iters = 0;
};

const int kSign2[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};

void helm_build(Helm &H) {
  const int d = H.dim, ne = H.ne;
  const int NX = H.nx + 1, NY = H.ny + 1;
  H.numnp = (i64)NX * NY * (d == 3 ? H.nz + 1 : 1);
  H.nelx = (i64)H.nx * H.ny * (d == 3 ? H.nz : 1);
  H.conn.assign(H.nelx * ne, 0);
  H.ke.assign(H.nelx * ne * ne, 0.0);
  H.nj.assign(H.nelx * H.nip * ne, 0.0);
  H.nq.assign(H.nip * ne, 0.0);
  H.diag.assign(H.numnp, 0.0);
  const double g = 1.0 / std::sqrt(3.0);
  double pts[8][3];
  if (d == 2) {
    const double p2[4][2] = {{-g, -g}, {g, -g}, {-g, g}, {g, g}};
    for (int ip = 0; ip < 4; ++ip) pts[ip][0] = p2[ip][0], pts[ip][1] = p2[ip][1], pts[ip][2] = 0;
  } else {
    for (int ip = 0; ip < 8; ++ip)
      for (int k = 0; k < 3; ++k) pts[ip][k] = kSign[ip][k] * g;
  }
  const double c = std::pow(H.radius / (2 * std::sqrt(3.0)), 2);
  for (i64 e = 0; e < H.nelx; ++e) {
    const i64 ex = e % H.nx, ey = (e / H.nx) % H.ny, ez = d == 3 ? e / ((i64)H.nx * H.ny) : 0;
    double Xe[8][3];
    for (int a = 0; a < ne; ++a) {
      const int sx = d == 2 ? kSign2[a][0] : kSign[a][0];
      const int sy = d == 2 ? kSign2[a][1] : kSign[a][1];
      const int sz = d == 2 ? -1 : kSign[a][2];
      const i64 ix = ex + (sx > 0), iy = ey + (sy > 0), iz = ez + (sz > 0);
      H.conn[e * ne + a] = (int)(ix + NX * (iy + (i64)NY * iz));
      Xe[a][0] = H.h[0] * ix;
      Xe[a][1] = H.h[1] * iy;
      Xe[a][2] = H.h[2] * iz;
    }
    double *K = &H.ke[e * ne * ne];
    for (int ip = 0; ip < H.nip; ++ip) {
      double N[8], dr[3][8];
      for (int a = 0; a < ne; ++a) {
        const double s[3] = {(double)(d == 2 ? kSign2[a][0] : kSign[a][0]),
                             (double)(d == 2 ? kSign2[a][1] : kSign[a][1]),
                             d == 2 ? 0.0 : (double)kSign[a][2]};
        double f[3];
        for (int k = 0; k < d; ++k) f[k] = 0.5 * (1.0 + s[k] * pts[ip][k]);
        N[a] = 1.0;
        for (int k = 0; k < d; ++k) N[a] *= f[k];
        for (int k = 0; k < d; ++k) {
          double v = 0.5 * s[k];
          for (int l = 0; l < d; ++l)
            if (l != k) v *= f[l];
          dr[k][a] = v;
        }
      }
      double J[3][3] = {{0}}, Ji[3][3] = {{0}}, det;
      for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
          double s = 0.0;
          for (int a = 0; a < ne; ++a) s += dr[i][a] * Xe[a][j];
          J[i][j] = s;
        }
      if (d == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        Ji[0][0] = J[1][1] / det;
        Ji[0][1] = -J[0][1] / det;
        Ji[1][0] = -J[1][0] / det;
        Ji[1][1] = J[0][0] / det;
      } else {
        det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
              J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
      }
      double dX[3][8];
      for (int i = 0; i < d; ++i)
        for (int a = 0; a < ne; ++a) {
          double s = 0.0;
          for (int k = 0; k < d; ++k) s += Ji[i][k] * dr[k][a];
          dX[i][a] = s;
        }
      const double jac = det * 1.0;
      for (int a = 0; a < ne; ++a) {
        for (int b = 0; b < ne; ++b) {
          double bb = 0.0;
          for (int k = 0; k < d; ++k) bb += dX[k][a] * dX[k][b];
          K[a * ne + b] += (N[a] * N[b] + c * bb) * jac;
        }
        H.nj[(e * H.nip + ip) * ne + a] = N[a] * jac;
        H.nq[ip * ne + a] = N[a];
      }
    }
    for (int a = 0; a < ne; ++a) H.diag[H.conn[e * ne + a]] += K[a * ne + a];
  }
}

// y = K x, element by element (the assembled globalK of filter_helm.hpp:117-118)
void helm_apply(const Helm &H, const double *x, double *y) {
  std::fill(y, y + H.numnp, 0.0);
  const int ne = H.ne;
  for (i64 e = 0; e < H.nelx; ++e) {
    const int *cn = &H.conn[e * ne];
    const double *K = &H.ke[e * ne * ne];
    for (int a = 0; a < ne; ++a) {
      double s = 0.0;
      for (int b = 0; b < ne; ++b) s += K[a * ne + b] * x[cn[b]];
      y[cn[a]] += s;
    }
  }
}

// Jacobi-preconditioned CG from zero, stop at ||r|| <= tol ||b|| (the
// ConjugateGradient + DiagonalPreconditioner of "Eigen_CG").  Returns the
// iteration count, -1 on the cap.
int helm_pcg(const Helm &H, const double *b, double *x, double tol, int maxit) {
  const i64 n = H.numnp;
  std::vector<double> r(b, b + n), z(n), p(n), q(n);
  std::fill(x, x + n, 0.0);
  const double bn2 = dot(b, b, (int)n);
  if (bn2 == 0.0) return 0;
  for (i64 i = 0; i < n; ++i) z[i] = r[i] / H.diag[i];
  p = z;
  double rho = dot(r.data(), z.data(), (int)n);
  for (int it = 1; it <= maxit; ++it) {
    helm_apply(H, p.data(), q.data());
    const double alpha = rho / dot(p.data(), q.data(), (int)n);
    for (i64 i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    if (dot(r.data(), r.data(), (int)n) <= tol * tol * bn2) return it;
    for (i64 i = 0; i < n; ++i) z[i] = r[i] / H.diag[i];
    const double rho1 = dot(r.data(), z.data(), (int)n);
    const double beta = rho1 / rho;
    rho = rho1;
    for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  return -1;
}

// FilterHelmholtz::densityFilter(design_s, "Eigen_CG", 0, sensFlag)
// (filter_helm.hpp:76-135).  Returns the CG iteration count (-1: cap).
int helm_filter(Helm &H, const double *in, double *out, int sens, double tol, int maxit) {
  const int ne = H.ne;
  std::vector<double> F(H.numnp, 0.0), val(H.numnp, 0.0);
  for (i64 e = 0; e < H.nelx; ++e) {
    double fe[8] = {0};
    for (int ip = 0; ip < H.nip; ++ip)
      for (int a = 0; a < ne; ++a) fe[a] += in[e] * H.nj[(e * H.nip + ip) * ne + a];
    for (int a = 0; a < ne; ++a) F[H.conn[e * ne + a]] += sens ? fe[a] / H.nip : fe[a];
  }
  const int it = helm_pcg(H, F.data(), val.data(), tol, maxit);
  for (i64 e = 0; e < H.nelx; ++e) {
    double s = 0.0;
    for (int ip = 0; ip < H.nip; ++ip) {
      double t = 0.0;
      for (int a = 0; a < ne; ++a) t += H.nq[ip * ne + a] * val[H.conn[e * ne + a]];
      s += t;
    }
    out[e] = sens ? s : s / H.nip;
  }
  H.last_iters = it;
  return it;
}

// ------------------------------------------------------------------- MMA --
// Svanberg's MMA with the dual interior-point subproblem solver, restated
// step for step from the published method and the reference's parameter
// choices (reference optimization/opt-solver/MMASolver.cpp:41-553).
struct MMA {
  int n, m, iter = 0;
  double xmamieps = 1.0e-5, epsimin, raa0 = 1.0e-5;
  double move = 0.15, albefa = 0.2, asyminit = 0.2, asymdec = 0.7, asyminc = 1.2;
  std::vector<double> a, c, d, y, lam, mu, s, low, upp, alpha, beta, p0, q0, pij,
      qij, b, grad, hess, xold1, xold2;
  double z = 0.0;
  MMA(int n_, int m_, double ai, double ci, double di)
      : n(n_), m(m_), epsimin(std::sqrt((double)(n_ + m_)) * 1e-9), a(m_, ai),
        c(m_, ci), d(m_, di), y(m_), lam(m_), mu(m_), s(2 * m_), low(n_), upp(n_),
        alpha(n_), beta(n_), p0(n_), q0(n_), pij((size_t)n_ * m_), qij((size_t)n_ * m_),
        b(m_), grad(m_), hess((size_t)m_ * m_), xold1(n_), xold2(n_) {}

  // Move asymptotes and build the convex separable approximation.
  void subproblem(const double *x, const double *df, const double *g,
                  const double *dg, const double *xmin, const double *xmax) {
    move_asymptotes(x, xmin, xmax);
    approximation(x, df, g, dg, xmin, xmax);
  }

  // Asymptotes L, U of this outer iteration (reference MMASolver.cpp:404-458).
  void move_asymptotes(const double *x, const double *xmin, const double *xmax) {
    ++iter;
    for (int i = 0; i < n; ++i) {
      double span = xmax[i] - xmin[i];
      if (iter < 3) {
        low[i] = x[i] - asyminit * span;
        upp[i] = x[i] + asyminit * span;
      } else {
        double osc = (x[i] - xold1[i]) * (xold1[i] - xold2[i]);
        double gam = osc < 0.0 ? asymdec : (osc > 0.0 ? asyminc : 1.0);
        low[i] = x[i] - gam * (xold1[i] - low[i]);
        upp[i] = x[i] + gam * (upp[i] - xold1[i]);
        double w = std::max(xmamieps, span);
        low[i] = std::min(std::max(low[i], x[i] - 100.0 * w), x[i] - 1.0e-5 * w);
        upp[i] = std::min(std::max(upp[i], x[i] + 1.0e-5 * w), x[i] + 100.0 * w);
        double lo = xmin[i] - 1.0e-6, hi = xmax[i] + 1.0e-6;
        if (x[i] < lo) {
          low[i] = x[i] - (hi - x[i]) / 0.9;
          upp[i] = x[i] + (hi - x[i]) / 0.9;
        }
        if (x[i] > hi) {
          low[i] = x[i] - (x[i] - lo) / 0.9;
          upp[i] = x[i] + (x[i] - lo) / 0.9;
        }
      }
    }
  }

  // Move limits and the approximation's p, q and b (MMASolver.cpp:460-512).
  void approximation(const double *x, const double *df, const double *g, const double *dg,
                     const double *xmin, const double *xmax) {
    for (int i = 0; i < n; ++i) {
      double span = xmax[i] - xmin[i];
      double al = std::max(xmin[i], low[i] + albefa * (x[i] - low[i]));
      al = std::max(al, x[i] - move * span);
      alpha[i] = std::min(al, xmax[i]);
      double be = std::min(xmax[i], upp[i] - albefa * (upp[i] - x[i]));
      be = std::min(be, x[i] + move * span);
      beta[i] = std::max(be, xmin[i]);
      double inv = 1.0 / std::max(xmamieps, span);
      double ux2 = std::pow(upp[i] - x[i], 2.0), xl2 = std::pow(x[i] - low[i], 2.0);
      double r0 = 0.001 * std::abs(df[i]) + raa0 * inv;
      p0[i] = ux2 * (std::max(0.0, df[i]) + r0);
      q0[i] = xl2 * (std::max(0.0, -1.0 * df[i]) + r0);
      for (int j = 0; j < m; ++j) {
        double gij = dg[(size_t)i * m + j];
        double rj = 0.001 * std::abs(gij) + raa0 * inv;
        pij[(size_t)i * m + j] = ux2 * (std::max(0.0, gij) + rj);
        qij[(size_t)i * m + j] = xl2 * (std::max(0.0, -1.0 * gij) + rj);
      }
    }
    for (int j = 0; j < m; ++j) {
      double acc = -g[j];
      for (int i = 0; i < n; ++i)
        acc += pij[(size_t)i * m + j] / (upp[i] - x[i]) + qij[(size_t)i * m + j] / (x[i] - low[i]);
      b[j] = acc;
    }
  }

  // Primal x(lambda), y(lambda), z(lambda) in closed form.
  void primal(double *x) {
    double la = 0.0;
    for (int j = 0; j < m; ++j) {
      if (lam[j] < 0.0) lam[j] = 0;
      y[j] = std::max(0.0, lam[j] - c[j]);
      la += lam[j] * a[j];
    }
    z = std::max(0.0, 10.0 * (la - 1.0));
    for (int i = 0; i < n; ++i) {
      double pl = p0[i], ql = q0[i];
      for (int j = 0; j < m; ++j) {
        pl += pij[(size_t)i * m + j] * lam[j];
        ql += qij[(size_t)i * m + j] * lam[j];
      }
      double xi = (std::sqrt(pl) * low[i] + std::sqrt(ql) * upp[i]) / (std::sqrt(pl) + std::sqrt(ql));
      xi = xi < alpha[i] ? alpha[i] : xi;
      x[i] = xi > beta[i] ? beta[i] : xi;
    }
  }

  double dual_sum(const double *x, int j) const {
    double acc = 0.0;
    for (int i = 0; i < n; ++i)
      acc += pij[(size_t)i * m + j] / (upp[i] - x[i]) + qij[(size_t)i * m + j] / (x[i] - low[i]);
    return acc;
  }

  void gradient(const double *x) {
    for (int j = 0; j < m; ++j) grad[j] = -b[j] - a[j] * z - y[j] + dual_sum(x, j);
  }

  void hessian(const double *x) {
    std::vector<double> df2(n), PQ((size_t)n * m);
    for (int i = 0; i < n; ++i) {
      double pl = p0[i], ql = q0[i];
      double ux = upp[i] - x[i], xl = x[i] - low[i];
      for (int j = 0; j < m; ++j) {
        pl += pij[(size_t)i * m + j] * lam[j];
        ql += qij[(size_t)i * m + j] * lam[j];
        PQ[(size_t)i * m + j] = pij[(size_t)i * m + j] / std::pow(ux, 2.0) -
                                qij[(size_t)i * m + j] / std::pow(xl, 2.0);
      }
      df2[i] = -1.0 / (2.0 * pl / std::pow(ux, 3.0) + 2.0 * ql / std::pow(xl, 3.0));
      double xp = (std::sqrt(pl) * low[i] + std::sqrt(ql) * upp[i]) / (std::sqrt(pl) + std::sqrt(ql));
      if (xp < alpha[i] || xp > beta[i]) df2[i] = 0.0;
    }
    for (int r = 0; r < m; ++r)
      for (int q = 0; q < m; ++q) {
        double acc = 0.0;
        for (int i = 0; i < n; ++i) acc += PQ[(size_t)i * m + r] * df2[i] * PQ[(size_t)i * m + q];
        hess[(size_t)r * m + q] = acc;
      }
    double la = 0.0;
    for (int j = 0; j < m; ++j) {
      if (lam[j] < 0.0) lam[j] = 0.0;
      la += lam[j] * a[j];
      if (lam[j] > c[j]) hess[(size_t)j * m + j] += -1.0;
      hess[(size_t)j * m + j] += -mu[j] / lam[j];
    }
    if (la > 0.0)
      for (int r = 0; r < m; ++r)
        for (int q = 0; q < m; ++q) hess[(size_t)r * m + q] += -10.0 * a[r] * a[q];
    double tr = 0.0;
    for (int j = 0; j < m; ++j) tr += hess[(size_t)j * m + j];
    double corr = 1e-4 * tr / m;
    if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
    for (int j = 0; j < m; ++j) hess[(size_t)j * m + j] += corr;
  }

  // Dense LU without pivoting and the two triangular solves.
  static void lu_solve(std::vector<double> &K, double *xv, int k) {
    for (int s0 = 0; s0 < k - 1; ++s0)
      for (int i = s0 + 1; i < k; ++i) {
        K[(size_t)i * k + s0] /= K[(size_t)s0 * k + s0];
        for (int j = s0 + 1; j < k; ++j) K[(size_t)i * k + j] -= K[(size_t)i * k + s0] * K[(size_t)s0 * k + j];
      }
    for (int i = 1; i < k; ++i) {
      double acc = 0.0;
      for (int j = 0; j < i; ++j) acc -= K[(size_t)i * k + j] * xv[j];
      xv[i] += acc;
    }
    xv[k - 1] /= K[(size_t)(k - 1) * k + (k - 1)];
    for (int i = k - 2; i >= 0; --i) {
      double acc = xv[i];
      for (int j = i + 1; j < k; ++j) acc -= K[(size_t)i * k + j] * xv[j];
      xv[i] = acc / K[(size_t)i * k + i];
    }
  }

  void line_search() {
    double th = 1.005;
    for (int j = 0; j < m; ++j) {
      th = std::max(th, -1.01 * s[j] / lam[j]);
      th = std::max(th, -1.01 * s[j + m] / mu[j]);
    }
    th = 1.0 / th;
    for (int j = 0; j < m; ++j) {
      lam[j] += th * s[j];
      mu[j] += th * s[j + m];
    }
  }

  double kkt_residual(const double *x, double epsi) const {
    double r = 0.0;
    for (int j = 0; j < m; ++j) {
      double r1 = -b[j] - a[j] * z - y[j] + mu[j] + dual_sum(x, j);
      double r2 = mu[j] * lam[j] - epsi;
      r = std::max(r, std::max(std::abs(r1), std::abs(r2)));
    }
    return r;
  }

  int solve_dual(double *x) {
    for (int j = 0; j < m; ++j) {
      lam[j] = c[j] / 2.0;
      mu[j] = 1.0;
    }
    double epsi = 1.0, err = 1.0;
    int outer = 0;
    while (epsi > epsimin && outer < 100) {
      ++outer;
      int inner = 0;
      while (err > 0.9 * epsi && inner < 100) {
        ++inner;
        primal(x);
        gradient(x);
        for (int j = 0; j < m; ++j) grad[j] = -1.0 * grad[j] - epsi / lam[j];
        hessian(x);
        if (m > 1) {
          lu_solve(hess, grad.data(), m);
          for (int j = 0; j < m; ++j) s[j] = grad[j];
        } else if (m > 0) {
          s[0] = grad[0] / hess[0];
        }
        for (int j = 0; j < m; ++j) s[m + j] = -mu[j] + epsi / lam[j] - s[j] * mu[j] / lam[j];
        line_search();
        primal(x);
        err = kkt_residual(x, epsi);
      }
      epsi *= 0.1;
    }
    return outer;
  }

  void update(double *x, const double *df, const double *g, const double *dg,
              const double *xmin, const double *xmax) {
    subproblem(x, df, g, dg, xmin, xmax);
    xold2 = xold1;
    std::copy(x, x + n, xold1.begin());
    solve_dual(x);
  }
};

// ----------------------------------------------------------------- GCMMA --
// Globally convergent MMA (K. Svanberg, A class of globally convergent
// optimization methods based on conservative convex separable approximations,
// SIAM J. Optim. 12, 2002) with the reference's choices (optimization/
// opt-solver/GCMMASolver.cpp:41-661): a = 0, c = 1000, d = 0; raa0 = raa_j = 1
// at construction, scaled by 0.1 (floor 1e-6) at every outer iteration and
// grown by the inner iterations until the approximations are conservative.
// The dual interior-point solver is the one of MMA above.
struct GCMMA : MMA {
  double raa0eps = 1.0e-6, raaeps = 1.0e-6;
  std::vector<double> raa, r, fapp;
  double r0 = 0.0, f0app = 0.0;
  GCMMA(int n_, int m_, double ai, double ci, double di)
      : MMA(n_, m_, ai, ci, di), raa(m_, 1.0), r(m_), fapp(m_) {
    raa0 = 1.0;
  }

  // Asymp (GCMMASolver.cpp:486-560): MMA's asymptotes, then raa0/raa shrink
  void asymptotes(const double *x, const double *xmin, const double *xmax) {
    move_asymptotes(x, xmin, xmax);
    raa0 = std::max(raa0eps, 0.1 * raa0);
    for (int j = 0; j < m; ++j) raa[j] = std::max(raaeps, 0.1 * raa[j]);
  }

  // GenSub (GCMMASolver.cpp:593-640): move limits without clamping to the
  // bounds again, raa0 in the objective and raa_j in constraint j, and the
  // constants r0, r_j that make the approximations exact at x.
  void gensub(const double *x, double f0x, const double *df0, const double *fx,
              const double *dfx, const double *xmin, const double *xmax) {
    r0 = 0.0;
    std::fill(r.begin(), r.end(), 0.0);
    for (int i = 0; i < n; ++i) {
      const double span = xmax[i] - xmin[i];
      alpha[i] = std::max(std::max(xmin[i], low[i] + albefa * (x[i] - low[i])),
                          x[i] - move * span);
      beta[i] = std::min(std::min(xmax[i], upp[i] - albefa * (upp[i] - x[i])),
                         x[i] + move * span);
      const double uxinv = 1.0 / (upp[i] - x[i]), xlinv = 1.0 / (x[i] - low[i]);
      const double inv = 1.0 / std::max(xmamieps, span);
      const double ux2 = std::pow(upp[i] - x[i], 2.0), xl2 = std::pow(x[i] - low[i], 2.0);
      const double pq0 = 0.001 * std::abs(df0[i]) + raa0 * inv;
      p0[i] = ux2 * (std::max(0.0, df0[i]) + pq0);
      q0[i] = xl2 * (std::max(0.0, -1.0 * df0[i]) + pq0);
      r0 += p0[i] * uxinv + q0[i] * xlinv;
      for (int j = 0; j < m; ++j) {
        const double gij = dfx[(size_t)i * m + j];
        const double pq = 0.001 * std::abs(gij) + raa[j] * inv;
        pij[(size_t)i * m + j] = ux2 * (std::max(0.0, gij) + pq);
        qij[(size_t)i * m + j] = xl2 * (std::max(0.0, -1.0 * gij) + pq);
        r[j] += pij[(size_t)i * m + j] * uxinv + qij[(size_t)i * m + j] * xlinv;
      }
    }
    r0 = f0x - r0;
    for (int j = 0; j < m; ++j) {
      r[j] = fx[j] - r[j];
      b[j] = -r[j];
    }
  }

  // ComputeApprox (GCMMASolver.cpp:642-661): the approximations at xmma
  void approx(const double *xm) {
    f0app = 0.0;
    std::fill(fapp.begin(), fapp.end(), 0.0);
    for (int i = 0; i < n; ++i) {
      const double uxinv = 1.0 / (upp[i] - xm[i]), xlinv = 1.0 / (xm[i] - low[i]);
      f0app += p0[i] * uxinv + q0[i] * xlinv;
      for (int j = 0; j < m; ++j)
        fapp[j] += pij[(size_t)i * m + j] * uxinv + qij[(size_t)i * m + j] * xlinv;
    }
    f0app += r0;
    for (int j = 0; j < m; ++j) fapp[j] += r[j];
  }

  // RaaUpdate (GCMMASolver.cpp:562-591): grow raa where the approximation
  // under-estimated the true value at xmma
  void raa_update(const double *xm, const double *xv, double f0new, const double *fnew,
                  const double *xmin, const double *xmax) {
    double cof = 0.0;
    for (int i = 0; i < n; ++i) {
      const double w = std::max(xmamieps, xmax[i] - xmin[i]);
      const double a1 = (xm[i] - xv[i]) / (upp[i] - xm[i]);
      const double a2 = (xm[i] - xv[i]) / (xm[i] - low[i]);
      cof += a1 * a2 * ((upp[i] - low[i]) / w);
    }
    cof = std::max(1e-12, cof);
    if (f0new > f0app + 0.5 * epsimin)
      raa0 = std::min(1.1 * (raa0 + (1.0 / cof) * (f0new - f0app)), 10.0 * raa0);
    for (int j = 0; j < m; ++j)
      if (fnew[j] > fapp[j] + 0.5 * epsimin)
        raa[j] = std::min(1.1 * (raa[j] + (1.0 / cof) * (fnew[j] - fapp[j])), 10.0 * raa[j]);
  }

  bool conservative(double f0new, const double *fnew) const {
    if (f0app + epsimin < f0new) return false;
    for (int j = 0; j < m; ++j)
      if (fapp[j] + epsimin < fnew[j]) return false;
    return true;
  }

  void outer_update(double *xm, const double *xv, double f0x, const double *df0,
                    const double *fx, const double *dfx, const double *xmin,
                    const double *xmax) {
    asymptotes(xv, xmin, xmax);
    gensub(xv, f0x, df0, fx, dfx, xmin, xmax);
    xold2 = xold1;
    std::copy(xv, xv + n, xold1.begin());
    std::copy(xv, xv + n, xm);
    solve_dual(xm);
    approx(xm);
  }

  void inner_update(double *xm, double f0new, const double *fnew, const double *xv, double f0x,
                    const double *df0, const double *fx, const double *dfx,
                    const double *xmin, const double *xmax) {
    raa_update(xm, xv, f0new, fnew, xmin, xmax);
    gensub(xv, f0x, df0, fx, dfx, xmin, xmax);
    std::copy(xv, xv + n, xm);
    solve_dual(xm);
    approx(xm);
  }
};

// ------------------------------------------------------------------- RNG ---
// splitmix64 -> double in [0,1): the project's input generator (no reference
// counterpart).
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// =============================================================== C API =====
extern "C" {

typedef struct orc_simp_opts {
  int n_iter;          // design iterations
  int optimizer;       // 0 = OC, 1 = MMA
  int use_projection;  // 0 = plain method (no tanh projection)
  double cg_tol;       // relative residual tolerance of Scaling_CG
  int cg_maxit;
  double xmin, xmax;
} orc_simp_opts;

typedef struct orc_simp_stats {
  double assembly_s, pcg_s, filter_s, sens_s, update_s, total_s;
  long long pcg_iters;
} orc_simp_stats;

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void *orc_create(int nx, int ny, int nz, double lx, double ly, double lz, double young1,
                 double young2, double poisson, double penalty) {
  Problem *P = new Problem();
  P->nx = nx; P->ny = ny; P->nz = nz;
  P->lx = lx; P->ly = ly; P->lz = lz;
  P->young1 = young1; P->young2 = young2; P->poisson = poisson; P->penalty = penalty;
  // counts (reference io/mesh_icarat.hpp:106-111)
  P->numnp = (i64)(nx + 1) * (ny + 1) * (nz + 1);
  P->neq = P->numnp * 3;
  P->nelx = (i64)nx * ny * nz;
  // coordinates (reference io/mesh_icarat.hpp:463-478)
  P->X.resize(P->numnp * 3);
  for (int i = 0; i < nz + 1; ++i)
    for (int j = 0; j < ny + 1; ++j)
      for (int k = 0; k < nx + 1; ++k) {
        i64 t = k + (i64)j * (nx + 1) + (i64)i * (nx + 1) * (ny + 1);
        P->X[3 * t] = ((double)lx / nx) * k;
        P->X[3 * t + 1] = ((double)ly / ny) * j;
        P->X[3 * t + 2] = ((double)lz / nz) * i;
      }
  // connectivity (reference io/mesh_icarat.hpp:480-511)
  P->conn.resize(P->nelx * 8);
  const i64 sx = nx + 1, sxy = (i64)(nx + 1) * (ny + 1);
  for (int i = 0; i < nz; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nx; ++k) {
        i64 e = (i64)nx * ny * i + (i64)nx * j + k;
        i64 base = sxy * i + sx * j + k;
        int *c = &P->conn[8 * e];
        c[0] = (int)base;
        c[1] = (int)(base + 1);
        c[2] = (int)(base + 1 + sx);
        c[3] = (int)(base + sx);
        c[4] = (int)(base + sxy);
        c[5] = (int)(base + 1 + sxy);
        c[6] = (int)(base + 1 + sx + sxy);
        c[7] = (int)(base + sx + sxy);
      }
  P->hasDir.assign(P->numnp, 0);
  P->dflag.assign(P->numnp * 6, 0);
  P->dval.assign(P->numnp * 6, 0.0);
  P->dof.assign(P->numnp * 3, -1);
  // KE0 from the element at the origin with young = 1
  double Xe[8][3];
  element_X(*P, 0, Xe);
  P->vol_e = element_ke(Xe, 1.0, poisson, P->KE0);
  return P;
}

void orc_destroy(void *h) { delete (Problem *)h; }
long long orc_numnp(void *h) { return ((Problem *)h)->numnp; }
long long orc_nelx(void *h) { return ((Problem *)h)->nelx; }
long long orc_neq(void *h) { return ((Problem *)h)->neq; }
int orc_numeq(void *h) { return ((Problem *)h)->numeq; }

void orc_coords(void *h, double *xyz) {
  Problem &P = *(Problem *)h;
  std::copy(P.X.begin(), P.X.end(), xyz);
}
void orc_connectivity(void *h, int *conn) {
  Problem &P = *(Problem *)h;
  std::copy(P.conn.begin(), P.conn.end(), conn);
}

// Dirichlet box, tolerance 1e-7, flags OR-ed and values overwritten per node
// (reference io/boundary_condition_icarat.hpp:72-138,
// io/boundary_condition_base.hpp:19-33).
void orc_add_dirichlet(void *h, const int *flag6, const double *val6, const double *range6) {
  Problem &P = *(Problem *)h;
  const double tol = 1.0e-7;
  double x0 = range6[0] - tol, x1 = range6[1] + tol, y0 = range6[2] - tol,
         y1 = range6[3] + tol, z0 = range6[4] - tol, z1 = range6[5] + tol;
  for (i64 n = 0; n < P.numnp; ++n) {
    const double *x = &P.X[3 * n];
    if (x[0] >= x0 && x[0] <= x1 && x[1] >= y0 && x[1] <= y1 && x[2] >= z0 && x[2] <= z1) {
      P.hasDir[n] = 1;
      for (int k = 0; k < 3; ++k) {
        if (P.dflag[6 * n + k] != 1) P.dflag[6 * n + k] = flag6[k];
        P.dval[6 * n + k] = val6[k];
      }
    }
  }
}

// Free DOFs first in node order, constrained DOFs after, in node order
// (reference io/boundary_condition_base.hpp:35-95).
int orc_make_dof(void *h) {
  Problem &P = *(Problem *)h;
  int counter = 0;
  for (i64 n = 0; n < P.numnp; ++n)
    for (int j = 0; j < 3; ++j) {
      if (!P.hasDir[n] || P.dflag[6 * n + j] == 0) P.dof[3 * n + j] = counter++;
    }
  int cc = counter;
  for (i64 n = 0; n < P.numnp; ++n)
    if (P.hasDir[n])
      for (int j = 0; j < 3; ++j)
        if (P.dflag[6 * n + j] != 0) P.dof[3 * n + j] = cc++;
  P.numeq = counter;
  P.fext.assign(counter, 0.0);
  return counter;
}

void orc_dof(void *h, int *dof) {
  Problem &P = *(Problem *)h;
  std::copy(P.dof.begin(), P.dof.end(), dof);
}

void orc_dirichlet_values(void *h, double *val /*numnp*3*/) {
  Problem &P = *(Problem *)h;
  for (i64 n = 0; n < P.numnp; ++n)
    for (int j = 0; j < 3; ++j) val[3 * n + j] = P.hasDir[n] ? P.dval[6 * n + j] : 0.0;
}

// Neumann load: a fully degenerate box is a point load on every node in it;
// otherwise a distributed load integrated on every element face with 4 nodes
// in the box (2x2 Gauss on quad4) (reference io/boundary_condition_icarat.hpp
// :212-410, io/boundary_condition_base.cpp:16-187, basis/nmatrix_surface.hpp:59-88).
void orc_add_neumann(void *h, const double *val3, const double *range6) {
  Problem &P = *(Problem *)h;
  const double tol = 1.0e-7;
  bool isDist = !(range6[0] == range6[1] && range6[2] == range6[3] && range6[4] == range6[5]);
  double x0 = range6[0] - tol, x1 = range6[1] + tol, y0 = range6[2] - tol,
         y1 = range6[3] + tol, z0 = range6[4] - tol, z1 = range6[5] + tol;
  auto inside = [&](i64 n) {
    const double *x = &P.X[3 * n];
    return x0 <= x[0] && x[0] <= x1 && y0 <= x[1] && x[1] <= y1 && z0 <= x[2] && x[2] <= z1;
  };
  if (!isDist) {
    for (i64 n = 0; n < P.numnp; ++n)
      if (inside(n))
        for (int d = 0; d < 3; ++d) {
          int q = P.dof[3 * n + d];
          if (q < P.numeq) P.fext[q] += val3[d];
        }
    return;
  }
  static const int faces[7][4] = {{0, 0, 0, 0}, {0, 1, 2, 3}, {3, 0, 4, 7}, {0, 1, 5, 4},
                                  {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
  const double g = 1.0 / std::sqrt(3.0);
  const double gp[4][2] = {{-g, -g}, {g, -g}, {-g, g}, {g, g}};
  for (i64 e = 0; e < P.nelx; ++e) {
    int hits[8], nh = 0;
    for (int k = 0; k < 8; ++k) {
      if (!inside(P.conn[8 * e + k])) continue;
      hits[nh++] = k;
      if (nh != 4) continue;
      // side from the local index of the 4th hit (boundary_condition_base.cpp:57-72)
      int side;
      if (k == 3) side = 1;
      else if (k == 5) side = 3;
      else if (k == 6) side = 4;
      else if (hits[2] - hits[1] == 1) side = 2;
      else if (hits[1] - hits[0] == 1) side = 5;
      else side = 6;
      double Xs[4][3];
      int sdof[12];
      for (int a = 0; a < 4; ++a) {
        i64 nd = P.conn[8 * e + faces[side][a]];
        for (int d = 0; d < 3; ++d) {
          Xs[a][d] = P.X[3 * nd + d];
          sdof[3 * a + d] = P.dof[3 * nd + d];
        }
      }
      double fe[12] = {0};
      for (int ip = 0; ip < 4; ++ip) {
        double r0 = gp[ip][0], r1 = gp[ip][1];
        double N[4] = {0.25 * (1.0 - r0) * (1.0 - r1), 0.25 * (1.0 + r0) * (1.0 - r1),
                       0.25 * (1.0 + r0) * (1.0 + r1), 0.25 * (1.0 - r0) * (1.0 + r1)};
        double dN[2][4] = {{-0.25 * (1.0 - r1), 0.25 * (1.0 - r1), 0.25 * (1.0 + r1), -0.25 * (1.0 + r1)},
                           {-0.25 * (1.0 - r0), -0.25 * (1.0 + r0), 0.25 * (1.0 + r0), 0.25 * (1.0 - r0)}};
        double t1[3], t2[3];
        for (int d = 0; d < 3; ++d) {
          t1[d] = dN[0][0] * Xs[0][d] + dN[0][1] * Xs[1][d] + dN[0][2] * Xs[2][d] + dN[0][3] * Xs[3][d];
          t2[d] = dN[1][0] * Xs[0][d] + dN[1][1] * Xs[1][d] + dN[1][2] * Xs[2][d] + dN[1][3] * Xs[3][d];
        }
        double cx = t1[1] * t2[2] - t1[2] * t2[1], cy = t1[2] * t2[0] - t1[0] * t2[2],
               cz = t1[0] * t2[1] - t1[1] * t2[0];
        double jac = std::sqrt(cx * cx + cy * cy + cz * cz) * 1.0;
        for (int a = 0; a < 4; ++a)
          for (int d = 0; d < 3; ++d) fe[3 * a + d] += N[a] * val3[d] * jac;
      }
      for (int t = 0; t < 12; ++t)
        if (sdof[t] < P.numeq) P.fext[sdof[t]] += fe[t];
    }
  }
}

void orc_fext(void *h, double *f) {
  Problem &P = *(Problem *)h;
  std::copy(P.fext.begin(), P.fext.end(), f);
}

void orc_ke0(void *h, double *ke) { std::memcpy(ke, ((Problem *)h)->KE0, sizeof(double) * 576); }

// Per-element Gauss Ke with the element's own coordinates and a given young.
double orc_element_ke(void *h, long long e, double young, double *ke) {
  Problem &P = *(Problem *)h;
  double Xe[8][3];
  element_X(P, e, Xe);
  return element_ke(Xe, young, P.poisson, ke);
}

// Stores the physical densities and assembles the free-DOF K(x).
void orc_set_density(void *h, const double *xphys) {
  Problem &P = *(Problem *)h;
  P.dens.assign(xphys, xphys + P.nelx);
  P.young.resize(P.nelx);
  for (i64 e = 0; e < P.nelx; ++e) P.young[e] = simp(P.young1, P.young2, xphys[e], P.penalty);
  if (!P.patternBuilt) build_pattern(P);
  assemble_values(P);
}

// Densities without the assembled K (meshes whose CSR would not fit, e.g.
// 256x128x128 with ~1e9 non-zeros): the element-by-element operators below
// need only the SIMP young.
void orc_set_density_mf(void *h, const double *xphys) {
  Problem &P = *(Problem *)h;
  P.dens.assign(xphys, xphys + P.nelx);
  P.young.resize(P.nelx);
  for (i64 e = 0; e < P.nelx; ++e) P.young[e] = simp(P.young1, P.young2, xphys[e], P.penalty);
  P.rowptr.clear();
  P.col.clear();
  P.val.clear();
  P.patternBuilt = false;
}

// K.u on numeq vectors gathered node by node: row (n, d) of the reference's
// assembled K (misc/assembling.hpp:18-39) is sum over the <= 8 elements
// around node n of young_e * KE0[rows of n] . u_e, u_e gathered with the
// Dirichlet values at constrained DOFs (misc/assembling.hpp:295-358).
// Parallel over nodes, each row summed in a fixed element order.
void orc_apply_k_nodal(void *h, const double *u, const double *dirval_nodes, double *y) {
  Problem &P = *(Problem *)h;
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const i64 NX = nx + 1, NY = ny + 1;
#pragma omp parallel for schedule(static)
  for (i64 n = 0; n < P.numnp; ++n) {
    const int i = (int)(n % NX), j = (int)((n / NX) % NY), k = (int)(n / (NX * NY));
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < 8; ++s) {
      const int ei = i - 1 + (s & 1), ej = j - 1 + ((s >> 1) & 1), ek = k - 1 + ((s >> 2) & 1);
      if (ei < 0 || ei >= nx || ej < 0 || ej >= ny || ek < 0 || ek >= nz) continue;
      const i64 e = ei + (i64)nx * (ej + (i64)ny * ek);
      int a = -1;
      double ue[24];
      for (int b = 0; b < 8; ++b) {
        const i64 nb = P.conn[8 * e + b];
        if (nb == n) a = b;
        for (int d = 0; d < 3; ++d) {
          const int q = P.dof[3 * nb + d];
          ue[3 * b + d] = q < P.numeq ? u[q] : (dirval_nodes ? dirval_nodes[3 * nb + d] : 0.0);
        }
      }
      const double E = P.young[e];
      for (int d = 0; d < 3; ++d) {
        double t = 0.0;
        for (int c = 0; c < 24; ++c) t += E * P.KE0[(3 * a + d) * 24 + c] * ue[c];
        acc[d] += t;
      }
    }
    for (int d = 0; d < 3; ++d) {
      const int q = P.dof[3 * n + d];
      if (q < P.numeq) y[q] = acc[d];
    }
  }
}

// ||fext - K u|| / ||fext|| with the node-by-node K.u (the true residual of
// analysis/linear.hpp:90-103 without an assembled matrix).
double orc_true_residual_nodal(void *h, const double *sol) {
  Problem &P = *(Problem *)h;
  std::vector<double> dv(P.numnp * 3), y(P.numeq);
  orc_dirichlet_values(h, dv.data());
  orc_apply_k_nodal(h, sol, dv.data(), y.data());
  double a = 0.0, b = 0.0;
  for (int i = 0; i < P.numeq; ++i) {
    const double r = P.fext[i] - y[i];
    a += r * r;
    b += P.fext[i] * P.fext[i];
  }
  return std::sqrt(a) / std::sqrt(b);
}

long long orc_nnz(void *h) { return ((Problem *)h)->rowptr.empty() ? 0 : ((Problem *)h)->rowptr.back(); }

void orc_spmv(void *h, const double *x, double *y) { spmv(*(Problem *)h, x, y); }

void orc_diag(void *h, double *dg) {
  Problem &P = *(Problem *)h;
  for (int r = 0; r < P.numeq; ++r) {
    double v = 0.0;
    for (i64 t = P.rowptr[r]; t < P.rowptr[r + 1]; ++t)
      if (P.col[t] == r) v = P.val[t];
    dg[r] = v;
  }
}

// Element-by-element K·u on numeq vectors: gather with Dirichlet values at
// constrained DOFs, Ke·u_e, scatter-add to free DOFs (reference
// misc/assembling.hpp:295-358, problem/elasticity.cpp:244-251).
void orc_apply_k_elementwise(void *h, const double *u, const double *dirval_nodes, double *y) {
  Problem &P = *(Problem *)h;
  std::fill(y, y + P.numeq, 0.0);
  for (i64 e = 0; e < P.nelx; ++e) {
    double ue[24], E = P.young[e];
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) {
        i64 n = P.conn[8 * e + a];
        int q = P.dof[3 * n + d];
        ue[3 * a + d] = q < P.numeq ? u[q] : (dirval_nodes ? dirval_nodes[3 * n + d] : 0.0);
      }
    for (int r = 0; r < 24; ++r) {
      int q = P.dof[3 * (i64)P.conn[8 * e + r / 3] + r % 3];
      if (q >= P.numeq) continue;
      double s = 0.0;
      for (int c = 0; c < 24; ++c) s += E * P.KE0[r * 24 + c] * ue[c];
      y[q] += s;
    }
  }
}

// Jacobi-preconditioned CG with x0 = 0, stop when ||r|| < tol*||b|| or
// iter > maxIter (reference analysis/solver/CG.cpp:56-100).
int orc_scaling_cg(void *h, const double *rhs, double *sol, double tol, int maxit, double *rnorm_out) {
  Problem &P = *(Problem *)h;
  const int n = P.numeq;
  std::vector<double> invd(n), r(rhs, rhs + n), p(n), q(n), z(n);
  orc_diag(h, invd.data());
  for (int i = 0; i < n; ++i) invd[i] = invd[i] != 0.0 ? 1.0 / invd[i] : 1.0;
  std::fill(sol, sol + n, 0.0);
  for (int i = 0; i < n; ++i) p[i] = invd[i] * r[i];
  double bnorm = std::sqrt(dot(rhs, rhs, n));
  double rho = dot(p.data(), r.data(), n);
  int iter = 0;
  double rnorm = 0.0;
  while (true) {
    ++iter;
    spmv(P, p.data(), q.data());
    double alpha = rho / dot(p.data(), q.data(), n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      sol[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = invd[i] * r[i];
    }
    double rho1 = dot(z.data(), r.data(), n);
    double beta = rho1 / rho;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) p[i] = beta * p[i] + z[i];
    rho = rho1;
    rnorm = std::sqrt(dot(r.data(), r.data(), n));
    if (rnorm < tol * bnorm) break;
    if (iter > maxit) break;
  }
  if (rnorm_out) *rnorm_out = rnorm;
  return iter;
}

// True relative residual ||fext - fint(u)|| / ||fext|| with fint assembled
// element by element (reference analysis/linear.hpp:90-103).
double orc_true_residual(void *h, const double *sol) {
  Problem &P = *(Problem *)h;
  std::vector<double> dv(P.numnp * 3), y(P.numeq);
  orc_dirichlet_values(h, dv.data());
  orc_apply_k_elementwise(h, sol, dv.data(), y.data());
  double a = 0.0, b = 0.0;
  for (int i = 0; i < P.numeq; ++i) {
    double r = P.fext[i] - y[i];
    a += r * r;
    b += P.fext[i] * P.fext[i];
  }
  return std::sqrt(a) / std::sqrt(b);
}

// Nodal values from the numeq solution; constrained DOFs take their
// Dirichlet value and |v| < 1e-15 is flushed to 0
// (reference analysis/analysis_utility.hpp:8-31).
void orc_node_values(void *h, const double *sol, double *nodeval) {
  Problem &P = *(Problem *)h;
  for (i64 n = 0; n < P.numnp; ++n)
    for (int j = 0; j < 3; ++j) {
      int q = P.dof[3 * n + j];
      double v = q >= P.numeq ? P.dval[6 * n + j] : sol[q];
      if (std::abs(v) < 1.0e-15) v = 0.0;
      nodeval[3 * n + j] = v;
    }
}

// Compliance c = u . fext (reference opt_macro_oc.cpp:110).
double orc_compliance(void *h, const double *sol) {
  Problem &P = *(Problem *)h;
  double c = 0.0;
  for (int i = 0; i < P.numeq; ++i) c += sol[i] * P.fext[i];
  return c;
}

// dfds = -(E'/E) u_e^T Ke u_e with Ke = E(x) KE0, dgds = volume / vnorm
// (reference sample/algorithm/opt-macro-oc/opt_macro_oc.cpp:182-210,
// opt-macro-mma-nondesign/opt_macro_mma_nondesign.cpp:188-218).
void orc_sensitivity(void *h, const double *nodeval, double vnorm, double *dfds, double *dgds) {
  Problem &P = *(Problem *)h;
  const double p = P.penalty, y1 = P.young1, y2 = P.young2;
#pragma omp parallel for schedule(static)
  for (i64 e = 0; e < P.nelx; ++e) {
    double dens = P.dens[e];
    double young = simp(y1, y2, dens, p);
    double dyoung = p * std::pow(dens, p - 1.0) * (y2 - y1);
    double E = dyoung / young;
    double ue[24];
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) ue[3 * a + d] = nodeval[3 * (i64)P.conn[8 * e + a] + d];
    double s = 0.0;
    for (int r = 0; r < 24; ++r) {
      double t = 0.0;
      for (int c = 0; c < 24; ++c) t += young * P.KE0[r * 24 + c] * ue[c];
      s += ue[r] * t;
    }
    dfds[e] = -E * s;
    dgds[e] = P.vol_e / vnorm;
  }
}

// Element-mean stress and von Mises stress (reference problem/elasticity.cpp:
// 166-209, Elastic::makeMisesStress): strain = sum_ip B u_e / ipmax, stress =
// sum_ip De(E(x_e)) B u_e / ipmax, mises = sqrt(stress^T V stress) with
// V = [1 -.5 -.5; -.5 1 -.5; -.5 -.5 1] (+) 3 I.  u_e from the nodal values
// (Dirichlet values included), De with the SIMP Young's modulus of the
// densities last set.  stress (may be NULL): nelx x 6 Voigt.
void orc_mises(void *h, const double *nodeval, double *mises, double *stress) {
  Problem &P = *(Problem *)h;
#pragma omp parallel for schedule(static)
  for (i64 e = 0; e < P.nelx; ++e) {
    double Xe[8][3], ue[24], D[6][6];
    element_X(P, e, Xe);
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) ue[3 * a + d] = nodeval[3 * (i64)P.conn[8 * e + a] + d];
    De3(simp(P.young1, P.young2, P.dens[e], P.penalty), P.poisson, D);
    const double fac = 1.0 / 8;
    double sig[6] = {0, 0, 0, 0, 0, 0};
    for (int ip = 0; ip < 8; ++ip) {
      double B[6][24], eps[6];
      element_B(Xe, ip, B);
      for (int i = 0; i < 6; ++i) {
        double t = 0.0;
        for (int c = 0; c < 24; ++c) t += B[i][c] * ue[c];
        eps[i] = t;
      }
      for (int i = 0; i < 6; ++i) {
        double t = 0.0;
        for (int k = 0; k < 6; ++k) t += D[i][k] * eps[k];
        sig[i] += t * fac;
      }
    }
    const double q = sig[0] * (sig[0] - 0.5 * sig[1] - 0.5 * sig[2]) +
                     sig[1] * (-0.5 * sig[0] + sig[1] - 0.5 * sig[2]) +
                     sig[2] * (-0.5 * sig[0] - 0.5 * sig[1] + sig[2]) +
                     3.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5]);
    mises[e] = std::sqrt(q);
    if (stress)
      for (int i = 0; i < 6; ++i) stress[6 * e + i] = sig[i];
  }
}

// ---------------------------------------------------------------- filter --
void *orc_filter_create(void *h, double radius, int brute) {
  Filter *F = new Filter();
  F->P = (Problem *)h;
  F->radius = radius;
  F->n = F->P->nelx;
  F->aa.assign(F->n, 0.0);
  if (brute) neighbours_brute(*F);
  else neighbours_structured(*F);
  // normaliser for the chain rule
  for (i64 i = 0; i < F->n; ++i) {
    double a = 0.0;
    for (i64 t = F->rowptr[i]; t < F->rowptr[i + 1]; ++t) a += F->radius - F->dis[t];
    F->aa[i] = a;
  }
  return F;
}
void orc_filter_destroy(void *f) { delete (Filter *)f; }
long long orc_filter_nnz(void *f) { return ((Filter *)f)->rowptr.back(); }
void orc_filter_lists(void *f, long long *rowptr, int *ids, double *dis) {
  Filter &F = *(Filter *)f;
  std::copy(F.rowptr.begin(), F.rowptr.end(), rowptr);
  std::copy(F.ids.begin(), F.ids.end(), ids);
  std::copy(F.dis.begin(), F.dis.end(), dis);
}
int orc_density_filter(void *f, const double *in, double *out, int isSens) {
  return density_filter(*(Filter *)f, in, out, isSens);
}

// Filter::sensitivityFilter(design_s, dfds) (reference optimization/filter.hpp:
// 76-104): aa = sum_j w_ij, bb = sum_j (w_ij x_j dfds_j) / (x_j V_j), out_i =
// (x_i V_i) / (x_i aa) * bb, neighbours in ascending element id, the reference's
// operation order.
void orc_sensitivity_filter(void *f, const double *x, const double *dfds, double *out) {
  Filter &F = *(Filter *)f;
  const double V = F.P->vol_e;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < F.n; ++i) {
    double aa = 0.0, bb = 0.0;
    for (i64 t = F.rowptr[i]; t < F.rowptr[i + 1]; ++t) {
      const int j = F.ids[t];
      const double w = F.radius - F.dis[t];
      aa += w;
      bb += (w * x[j] * dfds[j]) / (x[j] * V);
    }
    const double cc = (x[i] * V) / (x[i] * aa);
    out[i] = cc * bb;
  }
}
void orc_threshold(long long n, const double *x, double beta, double T, double *out) {
  threshold(n, x, beta, T, out);
}
void orc_threshold_deriv(long long n, const double *x, double beta, double T, double *out) {
  threshold_deriv(n, x, beta, T, out);
}

// OC bisection with the driver's volume callback (filter -> optional
// projection -> sum x_hat * vol) (reference
// optimization/opt-solver/OCSolver.hpp:20-79, opt_macro_oc.cpp:153-164).
// Returns the number of bisection steps; lambda of the last step in *lam.
}  // extern "C"

// OC with the drivers' volume callback: every bisection step filters x' with
// `filt` (density mode), projects it if asked and sums the volume
// (optimization/opt-solver/OCSolver.hpp:20-79, opt_macro_oc.cpp:153-164).
template <class Filt>
int oc_update_impl(i64 n, double vol_e, Filt filt, double *x, const double *dfdx, double g,
                   const double *dgdx, const double *xmin, const double *xmax,
                   int use_projection, double beta, double *lam_out) {
  const double iota = 0.5, movelimit = 0.15, lambdaeps = 1.0e-5;
  const int maxiter = 100;
  double l0 = 0.0, l1 = 1.0e4, lam = 0.0;
  std::vector<double> xn(n), xf(n), xt(n);
  int counter = 0;
  while ((l1 - l0) / (l1 + l0) > lambdaeps && maxiter > counter) {
    lam = 0.5 * (l1 + l0);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
      double v = std::pow(std::max(0.0, -dfdx[i] / (dgdx[i] * lam)), iota) * x[i];
      double lo = std::max(xmin[i], (1.0 - movelimit) * x[i]);
      double hi = std::min(xmax[i], (1.0 + movelimit) * x[i]);
      if (v < lo) v = lo;
      else if (v > hi) v = hi;
      xn[i] = v;
    }
    filt(xn.data(), xf.data());
    const double *xs = xf.data();
    if (use_projection) {
      threshold(n, xf.data(), beta, 0.5, xt.data());
      xs = xt.data();
    }
    double vol = 0.0;
    for (i64 i = 0; i < n; ++i) vol += xs[i] * vol_e;
    if (vol - g > 0.0) l0 = lam;
    else l1 = lam;
    ++counter;
  }
  std::copy(xn.begin(), xn.end(), x);
  if (lam_out) *lam_out = lam;
  return counter;
}

extern "C" {

int orc_oc_update(void *f, double *x, const double *dfdx, double g, const double *dgdx,
                  const double *xmin, const double *xmax, int use_projection, double beta,
                  double *lam_out) {
  Filter &F = *(Filter *)f;
  return oc_update_impl(
      F.n, F.P->vol_e, [&](const double *in, double *out) { density_filter(F, in, out, 0); }, x,
      dfdx, g, dgdx, xmin, xmax, use_projection, beta, lam_out);
}

void *orc_mma_create(int n, int m, double a, double c, double d) { return new MMA(n, m, a, c, d); }
void orc_mma_destroy(void *h) { delete (MMA *)h; }
void orc_mma_set(void *h, double move, double albefa, double asyminit, double asymdec,
                 double asyminc) {
  MMA &M = *(MMA *)h;
  M.move = move; M.albefa = albefa; M.asyminit = asyminit; M.asymdec = asymdec; M.asyminc = asyminc;
}
void orc_mma_update(void *h, double *x, const double *dfdx, const double *gx, const double *dgdx,
                    const double *xmin, const double *xmax) {
  ((MMA *)h)->update(x, dfdx, gx, dgdx, xmin, xmax);
}
void orc_mma_state(void *h, double *low, double *upp) {
  MMA &M = *(MMA *)h;
  std::copy(M.low.begin(), M.low.end(), low);
  std::copy(M.upp.begin(), M.upp.end(), upp);
}

void *orc_gcmma_create(int n, int m, double a, double c, double d) {
  return new GCMMA(n, m, a, c, d);
}
void orc_gcmma_destroy(void *h) { delete (GCMMA *)h; }
void orc_gcmma_set(void *h, double move, double albefa, double asyminit, double asymdec,
                   double asyminc) {
  GCMMA &G = *(GCMMA *)h;
  G.move = move; G.albefa = albefa; G.asyminit = asyminit; G.asymdec = asymdec; G.asyminc = asyminc;
}
// GCMMASolver::OuterUpdate / InnerUpdate / ConCheck (GCMMASolver.hpp:69-85)
void orc_gcmma_outer(void *h, double *xmma, const double *xval, double f0x, const double *df0dx,
                     const double *fx, const double *dfdx, const double *xmin,
                     const double *xmax) {
  ((GCMMA *)h)->outer_update(xmma, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
}
void orc_gcmma_inner(void *h, double *xmma, double f0xnew, const double *fxnew,
                     const double *xval, double f0x, const double *df0dx, const double *fx,
                     const double *dfdx, const double *xmin, const double *xmax) {
  ((GCMMA *)h)->inner_update(xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
}
int orc_gcmma_concheck(void *h, double f0xnew, const double *fxnew) {
  return ((GCMMA *)h)->conservative(f0xnew, fxnew) ? 1 : 0;
}
// {raa0, raa[0..m), f0app, fapp[0..m)}
void orc_gcmma_state(void *h, double *out) {
  const GCMMA &G = *(GCMMA *)h;
  out[0] = G.raa0;
  for (int j = 0; j < G.m; ++j) out[1 + j] = G.raa[j];
  out[1 + G.m] = G.f0app;
  for (int j = 0; j < G.m; ++j) out[2 + G.m + j] = G.fapp[j];
}

// The GCMMA driver's loop (sample/algorithm/opt-macro-gcmma/opt_macro_gcmma.cpp:
// 84-212): filter -> projection with beta = min(int(max((k - 100) / 40, 1)), 8)
// (none when use_projection = 0) -> solve -> f = compliance, g = volume - V0 ->
// sensitivities with the projection derivative -> chain rule -> OuterUpdate;
// then the new design is evaluated (filter, projection, solve) and inner
// iterations (at most 15) run until the approximations are conservative.
// hist per iteration: {compliance, volume, PCG iterations of all solves,
// inner iterations}.
int orc_simp_run_gcmma(void *h, void *f, const orc_simp_opts *o, double *x, double *hist,
                       orc_simp_stats *st) {
  Problem &P = *(Problem *)h;
  Filter &F = *(Filter *)f;
  const i64 n = P.nelx;
  orc_simp_stats S;
  std::memset(&S, 0, sizeof(S));
  const double t_all = now_s();
  std::vector<double> xf(n), xt(n), thd(n, 1.0), dfds(n), dgds(n), tmp(n), xnew(n);
  std::vector<double> u(P.numeq), nodeval(P.numnp * 3);
  std::vector<double> xmin(n, o->xmin), xmax(n, o->xmax);
  GCMMA G((int)n, 1, 0.0, 1000.0, 0.0);
  double V0 = 0.0;
  long long its = 0;
  // filter, projection, solve at design xd; returns compliance, sets vol
  auto evaluate = [&](const double *xd, double beta, double &vol) {
    density_filter(F, xd, xf.data(), 0);
    const double *xp = xf.data();
    if (o->use_projection) {
      threshold(n, xf.data(), beta, 0.5, xt.data());
      threshold_deriv(n, xf.data(), beta, 0.5, thd.data());
      xp = xt.data();
    }
    orc_set_density(h, xp);
    double rn = 0.0;
    const int it = orc_scaling_cg(h, P.fext.data(), u.data(), o->cg_tol, o->cg_maxit, &rn);
    its += it;
    S.pcg_iters += it;
    vol = 0.0;
    for (i64 e = 0; e < n; ++e) vol += xp[e] * P.vol_e;
    return orc_compliance(h, u.data());
  };
  for (int step = 1; step <= o->n_iter; ++step) {
    its = 0;
    const int bt = (int)std::max((step - 100) / 40.0, 1.0);
    const double beta = std::min((double)bt, 8.0);
    double vol = 0.0;
    const double c = evaluate(x, beta, vol);
    if (step == 1) V0 = vol;
    const double g = vol - V0;
    orc_node_values(h, u.data(), nodeval.data());
    orc_sensitivity(h, nodeval.data(), 1.0, dfds.data(), dgds.data());
    for (i64 e = 0; e < n; ++e) {
      dfds[e] *= thd[e];
      dgds[e] *= thd[e];
    }
    density_filter(F, dfds.data(), tmp.data(), 1);
    dfds.swap(tmp);
    density_filter(F, dgds.data(), tmp.data(), 1);
    dgds.swap(tmp);
    G.outer_update(xnew.data(), x, c, dfds.data(), &g, dgds.data(), xmin.data(), xmax.data());
    double vnew = 0.0;
    double cnew = evaluate(xnew.data(), beta, vnew);
    double gnew = vnew - V0;
    bool ok = G.conservative(cnew, &gnew);
    int inner = 0;
    for (; !ok && inner < 15; ++inner) {
      G.inner_update(xnew.data(), cnew, &gnew, x, c, dfds.data(), &g, dgds.data(), xmin.data(),
                     xmax.data());
      cnew = evaluate(xnew.data(), beta, vnew);
      gnew = vnew - V0;
      ok = G.conservative(cnew, &gnew);
    }
    std::copy(xnew.begin(), xnew.end(), x);
    if (hist) {
      hist[4 * (step - 1) + 0] = c;
      hist[4 * (step - 1) + 1] = vol;
      hist[4 * (step - 1) + 2] = (double)its;
      hist[4 * (step - 1) + 3] = inner;
    }
  }
  S.total_s = now_s() - t_all;
  if (st) *st = S;
  return 0;
}

// One SIMP design loop as the reference drivers run it (filter -> optional
// projection -> solve -> compliance -> volume -> sensitivity -> chain rule ->
// OC or MMA), for n_iter iterations (reference sample/algorithm/opt-macro-oc/
// opt_macro_oc.cpp:91-178, opt-macro-mma-nondesign/opt_macro_mma_nondesign.cpp
// :90-178).  x holds the design variables on entry and exit.  hist receives
// per iteration {compliance, volume, pcg iterations, true residual}.
}  // extern "C"

// filt(in, out, isSens) is the density filter of the loop (cone or Helmholtz)
template <class Filt>
int simp_run_impl(void *h, Filt filt, const orc_simp_opts *o, double *x, double *hist,
                  orc_simp_stats *st) {
  Problem &P = *(Problem *)h;
  const i64 n = P.nelx;
  orc_simp_stats S;
  std::memset(&S, 0, sizeof(S));
  double t_all = now_s();
  std::vector<double> xf(n), xt(n), thd(n), dfds(n), dgds(n), tmp(n);
  std::vector<double> u(P.numeq), nodeval(P.numnp * 3);
  std::vector<double> xmin(n, o->xmin), xmax(n, o->xmax);
  MMA *mma = o->optimizer == 1 ? new MMA((int)n, 1, 0.0, 1e5, 0.0) : nullptr;
  double V0 = 0.0;
  for (int step = 1; step <= o->n_iter; ++step) {
    double t0 = now_s();
    filt(x, xf.data(), 0);
    double beta = std::min(std::pow(2, (int)(step / 40)), 16.0);
    const double *xp = xf.data();
    if (o->use_projection) {
      threshold(n, xf.data(), beta, 0.5, xt.data());
      threshold_deriv(n, xf.data(), beta, 0.5, thd.data());
      xp = xt.data();
    }
    double t1 = now_s();
    S.filter_s += t1 - t0;
    orc_set_density(h, xp);
    double t2 = now_s();
    S.assembly_s += t2 - t1;
    double rn = 0.0;
    int it = orc_scaling_cg(h, P.fext.data(), u.data(), o->cg_tol, o->cg_maxit, &rn);
    double t3 = now_s();
    S.pcg_s += t3 - t2;
    S.pcg_iters += it;
    double c = orc_compliance(h, u.data());
    double vol = 0.0;
    for (i64 e = 0; e < n; ++e) vol += xp[e] * P.vol_e;
    if (step == 1) V0 = vol;
    orc_node_values(h, u.data(), nodeval.data());
    orc_sensitivity(h, nodeval.data(), o->optimizer == 1 ? V0 : 1.0, dfds.data(), dgds.data());
    if (o->use_projection)
      for (i64 e = 0; e < n; ++e) {
        dfds[e] *= thd[e];
        dgds[e] *= thd[e];
      }
    double t4 = now_s();
    S.sens_s += t4 - t3;
    filt(dfds.data(), tmp.data(), 1);
    dfds.swap(tmp);
    filt(dgds.data(), tmp.data(), 1);
    dgds.swap(tmp);
    double t5 = now_s();
    S.filter_s += t5 - t4;
    if (hist) {
      hist[4 * (step - 1) + 0] = c;
      hist[4 * (step - 1) + 1] = vol;
      hist[4 * (step - 1) + 2] = it;
      hist[4 * (step - 1) + 3] = rn;
    }
    if (o->optimizer == 1) {
      double g = step == 1 ? 0.0 : vol / V0 - 1.0;
      mma->update(x, dfds.data(), &g, dgds.data(), xmin.data(), xmax.data());
    } else {
      oc_update_impl(
          n, P.vol_e, [&](const double *in, double *out) { filt(in, out, 0); }, x, dfds.data(),
          V0, dgds.data(), xmin.data(), xmax.data(), o->use_projection, beta, nullptr);
    }
    S.update_s += now_s() - t5;
  }
  delete mma;
  S.total_s = now_s() - t_all;
  if (st) *st = S;
  return 0;
}

extern "C" {

int orc_simp_run(void *h, void *f, const orc_simp_opts *o, double *x, double *hist,
                 orc_simp_stats *st) {
  Filter &F = *(Filter *)f;
  return simp_run_impl(
      h, [&](const double *in, double *out, int sens) { density_filter(F, in, out, sens); }, o, x,
      hist, st);
}

// ---------------------------------------------------------- Helmholtz ----
// FilterHelmholtz<E,N>(fem, element, node, radius) on a structured grid of
// quad4 (dim 2) or hexa8 (dim 3) elements of size hx x hy (x hz)
// (optimization/filter_helm.hpp:69-75).
void *orc_helm_create(int dim, int nx, int ny, int nz, double hx, double hy, double hz,
                      double radius) {
  Helm *H = new Helm;
  H->dim = dim;
  H->nx = nx;
  H->ny = ny;
  H->nz = dim == 3 ? nz : 1;
  H->h[0] = hx;
  H->h[1] = hy;
  H->h[2] = hz;
  H->radius = radius;
  H->ne = H->nip = dim == 2 ? 4 : 8;
  helm_build(*H);
  return H;
}
void orc_helm_destroy(void *h) { delete (Helm *)h; }
long long orc_helm_numnp(void *h) { return ((Helm *)h)->numnp; }
// element matrix Ke (ne x ne) of element e, and its connectivity
void orc_helm_element(void *h, long long e, double *ke, int *conn) {
  const Helm &H = *(Helm *)h;
  std::copy(&H.ke[e * H.ne * H.ne], &H.ke[(e + 1) * H.ne * H.ne], ke);
  std::copy(&H.conn[e * H.ne], &H.conn[(e + 1) * H.ne], conn);
}
// y = K x (the assembled Helmholtz matrix, numnp)
void orc_helm_apply(void *h, const double *x, double *y) { helm_apply(*(Helm *)h, x, y); }
// densityFilter(in, "Eigen_CG", 0, sens) with CG tolerance tol; returns the
// CG iterations (-1 at the cap)
int orc_helm_filter(void *h, const double *in, double *out, int sens, double tol, int maxit) {
  return helm_filter(*(Helm *)h, in, out, sens, tol, maxit);
}
// OC whose volume callback filters with the Helmholtz filter
int orc_helm_oc_update(void *h, void *hp, double *x, const double *dfdx, double g,
                       const double *dgdx, const double *xmin, const double *xmax,
                       int use_projection, double beta, double tol) {
  Helm &H = *(Helm *)h;
  const Problem &P = *(Problem *)hp;
  return oc_update_impl(
      H.nelx, P.vol_e, [&](const double *in, double *out) { helm_filter(H, in, out, 0, tol, 100000); },
      x, dfdx, g, dgdx, xmin, xmax, use_projection, beta, nullptr);
}
// The SIMP loop with the Helmholtz filter in place of the cone filter
// (sample/algorithm/opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:84-150).
int orc_simp_run_helm(void *h, void *helm, const orc_simp_opts *o, double tol, double *x,
                      double *hist, orc_simp_stats *st) {
  Helm &H = *(Helm *)helm;
  return simp_run_impl(
      h, [&](const double *in, double *out, int sens) { helm_filter(H, in, out, sens, tol, 100000); },
      o, x, hist, st);
}

// Seeded smooth density field: sum of nblob isotropic Gaussians (amplitude
// U[0.5,1.5], centre U(domain), sigma U[smin,smax]) at element centres,
// rescaled to the target mean and clipped to [lo, hi].  Returns the mean
// after clipping.
double orc_gaussian_field(int nx, int ny, int nz, double h, unsigned long long seed, int nblob,
                          double smin, double smax, double mean, double lo, double hi, double *out) {
  SplitMix rng{seed};
  std::vector<double> A(nblob), MX(nblob), MY(nblob), MZ(nblob), S2(nblob);
  for (int k = 0; k < nblob; ++k) {
    A[k] = 0.5 + rng.uniform();
    MX[k] = rng.uniform() * nx * h;
    MY[k] = rng.uniform() * ny * h;
    MZ[k] = rng.uniform() * nz * h;
    double s = smin + (smax - smin) * rng.uniform();
    S2[k] = 2.0 * s * s;
  }
  const i64 n = (i64)nx * ny * nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        double cx = (i + 0.5) * h, cy = (j + 0.5) * h, cz = (k + 0.5) * h;
        double g = 0.0;
        for (int b = 0; b < nblob; ++b) {
          double dx = cx - MX[b], dy = cy - MY[b], dz = cz - MZ[b];
          g += A[b] * std::exp(-(dx * dx + dy * dy + dz * dz) / S2[b]);
        }
        out[i + (i64)nx * (j + (i64)ny * k)] = g;
      }
  double sum = 0.0;
  for (i64 e = 0; e < n; ++e) sum += out[e];
  double scale = mean / (sum / (double)n);
  double s2 = 0.0;
  for (i64 e = 0; e < n; ++e) {
    double v = out[e] * scale;
    v = v < lo ? lo : (v > hi ? hi : v);
    out[e] = v;
    s2 += v;
  }
  return s2 / (double)n;
}

}  // extern "C"
