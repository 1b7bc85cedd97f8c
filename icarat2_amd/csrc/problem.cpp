// problem.cpp -- host-side problem setup (see problem.hpp).
#include "problem.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "hip_util.hpp"

namespace icarat_hip {

namespace {

// natural coordinates of hexa8 corner a (reference basis/shape_function.cpp:403-447)
inline int corner_sign(int a, int d) {
  const int o = d == 0 ? lnode_x(a) : (d == 1 ? lnode_y(a) : lnode_z(a));
  return o ? 1 : -1;
}

}  // namespace

double hexa8_stiffness(double hx, double hy, double hz, double nu, double ke[576]) {
  // The 2x2x2 rule integrates the trilinear element exactly; evaluated in
  // extended precision and rounded once, every entry is the FP64 rounding of
  // the exact integral, so entries equal in exact arithmetic (the symmetry
  // orbits the stencil kernels rely on) are bit-identical.
  using R = long double;
  const R lam = (R)nu / ((1 + (R)nu) * (1 - 2 * (R)nu));
  const R mu = 1 / (2 * (1 + (R)nu));
  const R g = 1 / std::sqrt((R)3);
  const R h[3] = {hx, hy, hz};
  R acc[576];
  for (int i = 0; i < 576; ++i) acc[i] = 0;
  R vol = 0;
  // Gauss points in the reference order (basis/gauss_integration.cpp:268-289)
  for (int ip = 0; ip < 8; ++ip) {
    const R r[3] = {corner_sign(ip, 0) * g, corner_sign(ip, 1) * g, corner_sign(ip, 2) * g};
    // Jacobian of the map from the unit cube [-1,1]^3 to the element
    R J[3] = {0, 0, 0};
    R dNr[8][3];
    for (int a = 0; a < 8; ++a) {
      const R s[3] = {(R)corner_sign(a, 0), (R)corner_sign(a, 1), (R)corner_sign(a, 2)};
      dNr[a][0] = s[0] * (1 + s[1] * r[1]) * (1 + s[2] * r[2]) / 8;
      dNr[a][1] = s[1] * (1 + s[2] * r[2]) * (1 + s[0] * r[0]) / 8;
      dNr[a][2] = s[2] * (1 + s[0] * r[0]) * (1 + s[1] * r[1]) / 8;
      for (int d = 0; d < 3; ++d) J[d] += dNr[a][d] * (lnode_x(a) * (d == 0) + lnode_y(a) * (d == 1) + lnode_z(a) * (d == 2)) * h[d];
    }
    const R detw = J[0] * J[1] * J[2];
    R gr[8][3];
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) gr[a][d] = dNr[a][d] / J[d];
    // K_{ai,bj} += w detJ (lam g_ai g_bj + mu (g_aj g_bi + delta_ij g_a.g_b))
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        const R gg = gr[a][0] * gr[b][0] + gr[a][1] * gr[b][1] + gr[a][2] * gr[b][2];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            R v = lam * gr[a][i] * gr[b][j] + mu * gr[a][j] * gr[b][i];
            if (i == j) v += mu * gg;
            acc[(3 * a + i) * 24 + 3 * b + j] += v * detw;
          }
      }
    vol += detw;
  }
  // an entry that vanishes exactly comes out at the rounding level of the
  // extended sums (1e-19 of the largest); every other exact entry is a ratio
  // of small integers times lam, mu and the element sizes, far above 1e-12
  R amax = 0;
  for (int i = 0; i < 576; ++i) amax = std::max(amax, std::fabs(acc[i]));
  for (int i = 0; i < 576; ++i) ke[i] = std::fabs(acc[i]) < 1e-12L * amax ? 0.0 : (double)acc[i];
  return (double)vol;
}

Problem::Problem(const GridSpec &g) : g_(g) {
  if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0) throw Error(kInvalidArgument, "grid divisions must be positive");
  if (!(g.lx > 0 && g.ly > 0 && g.lz > 0)) throw Error(kInvalidArgument, "grid lengths must be positive");
  if (!(g.young2 - g.young1 >= 0.0)) throw Error(kInvalidArgument, "young2 must be >= young1");
  const long long nmax = 2147483647LL;
  L_.nx = g.nx; L_.ny = g.ny; L_.nz = g.nz;
  L_.NX = g.nx + 1; L_.NY = g.ny + 1; L_.NZ = g.nz + 1;
  L_.numnp = (long long)L_.NX * L_.NY * L_.NZ;
  L_.nelx = (long long)g.nx * g.ny * g.nz;
  if (3 * L_.numnp > nmax) throw Error(kInvalidArgument, "mesh too large for 32-bit DOF numbers");
  L_.Np = (L_.numnp + 255) / 256 * 256;
  dirflag_.assign(L_.numnp, 0);
  hasdir_.assign(L_.numnp, 0);
  dirval_.assign(L_.numnp * 3, 0.0);
  vol_e_ = hexa8_stiffness(hx(), hy(), hz(), g.poisson, ke0_);
}

double Problem::coord(long long n, int d) const {
  // node coordinate = (length / division) * index (reference io/mesh_icarat.hpp:463-478)
  if (d == 0) return ((double)g_.lx / g_.nx) * (double)(n % L_.NX);
  if (d == 1) return ((double)g_.ly / g_.ny) * (double)((n / L_.NX) % L_.NY);
  return ((double)g_.lz / g_.nz) * (double)(n / ((long long)L_.NX * L_.NY));
}

void Problem::add_dirichlet(const int flag[6], const double val[6], const double range[6]) {
  if (numeq_ >= 0) throw Error(kNotReady, "Dirichlet conditions must precede make_dof");
  {
    DirichletBox b;
    for (int i = 0; i < 6; ++i) {
      b.flag[i] = flag[i];
      b.val[i] = val[i];
      b.range[i] = range[i];
    }
    boxes_.push_back(b);
  }
  const double tol = 1.0e-7;
  const double lo[3] = {range[0] - tol, range[2] - tol, range[4] - tol};
  const double hi[3] = {range[1] + tol, range[3] + tol, range[5] + tol};
  for (long long n = 0; n < L_.numnp; ++n) {
    bool in = true;
    for (int d = 0; d < 3 && in; ++d) {
      const double x = coord(n, d);
      in = x >= lo[d] && x <= hi[d];
    }
    if (!in) continue;
    hasdir_[n] = 1;
    for (int c = 0; c < 3; ++c) {
      // a component once constrained stays constrained; the value is overwritten
      if (flag[c] != 0) dirflag_[n] |= (uint8_t)(1u << c);
      dirval_[3 * n + c] = val[c];
    }
  }
}

int Problem::make_dof() {
  dof_.assign(L_.numnp * 3, -1);
  free_.assign(L_.numnp, 0);
  int q = 0;
  for (long long n = 0; n < L_.numnp; ++n)
    for (int c = 0; c < 3; ++c)
      if (!((dirflag_[n] >> c) & 1u)) {
        dof_[3 * n + c] = q++;
        free_[n] |= (uint8_t)(1u << c);
      }
  numeq_ = q;
  for (long long n = 0; n < L_.numnp; ++n)
    for (int c = 0; c < 3; ++c)
      if ((dirflag_[n] >> c) & 1u) dof_[3 * n + c] = q++;
  cases_.assign(1, std::vector<double>(numeq_, 0.0));
  cur_ = 0;
  return numeq_;
}

int Problem::new_load_case() {
  if (numeq_ < 0) throw Error(kNotReady, "make_dof must precede load cases");
  cases_.emplace_back(numeq_, 0.0);
  cur_ = (int)cases_.size() - 1;
  return cur_;
}

void Problem::select_load_case(int i) {
  if (i < 0 || i >= (int)cases_.size()) throw Error(kInvalidArgument, "no such load case");
  cur_ = i;
}

bool Problem::has_nonzero_dirichlet() const {
  for (long long n = 0; n < L_.numnp; ++n)
    for (int c = 0; c < 3; ++c)
      if (((dirflag_[n] >> c) & 1u) && dirval_[3 * n + c] != 0.0) return true;
  return false;
}

void Problem::add_neumann(const double val[3], const double range[6]) {
  if (numeq_ < 0) throw Error(kNotReady, "make_dof must precede Neumann conditions");
  const double tol = 1.0e-7;
  const bool point = range[0] == range[1] && range[2] == range[3] && range[4] == range[5];
  const double lo[3] = {range[0] - tol, range[2] - tol, range[4] - tol};
  const double hi[3] = {range[1] + tol, range[3] + tol, range[5] + tol};
  auto inside = [&](long long n) {
    for (int d = 0; d < 3; ++d) {
      const double x = coord(n, d);
      if (!(lo[d] <= x && x <= hi[d])) return false;
    }
    return true;
  };
  if (point) {
    for (long long n = 0; n < L_.numnp; ++n)
      if (inside(n))
        for (int c = 0; c < 3; ++c) {
          const int q = dof_[3 * n + c];
          if (q < numeq_) fext()[q] += val[c];
        }
    return;
  }
  // face sides by local node index (reference io/boundary_condition_base.cpp:308-323)
  static const int side_nodes[7][4] = {{0, 0, 0, 0}, {0, 1, 2, 3}, {3, 0, 4, 7}, {0, 1, 5, 4},
                                       {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
  const double g = 1.0 / std::sqrt(3.0);
  const double gp[4][2] = {{-g, -g}, {g, -g}, {-g, g}, {g, g}};
  for (int k = 0; k < g_.nz; ++k)
    for (int j = 0; j < g_.ny; ++j)
      for (int i = 0; i < g_.nx; ++i) {
        const long long n0 = i + (long long)L_.NX * (j + (long long)L_.NY * k);
        int hits[8], nh = 0;
        for (int a = 0; a < 8; ++a) {
          const long long n = n0 + lnode_x(a) + (long long)L_.NX * (lnode_y(a) + (long long)L_.NY * lnode_z(a));
          if (!inside(n)) continue;
          hits[nh++] = a;
          if (nh != 4) continue;
          // side from the 4th hit (reference io/boundary_condition_base.cpp:57-72)
          int side = a == 3 ? 1 : a == 5 ? 3 : a == 6 ? 4
                   : (hits[2] - hits[1] == 1 ? 2 : (hits[1] - hits[0] == 1 ? 5 : 6));
          double X[4][3];
          long long nd[4];
          for (int s = 0; s < 4; ++s) {
            const int b = side_nodes[side][s];
            nd[s] = n0 + lnode_x(b) + (long long)L_.NX * (lnode_y(b) + (long long)L_.NY * lnode_z(b));
            for (int d = 0; d < 3; ++d) X[s][d] = coord(nd[s], d);
          }
          double fe[4][3] = {};
          for (int ip = 0; ip < 4; ++ip) {
            const double r0 = gp[ip][0], r1 = gp[ip][1];
            const double N[4] = {0.25 * (1.0 - r0) * (1.0 - r1), 0.25 * (1.0 + r0) * (1.0 - r1),
                                 0.25 * (1.0 + r0) * (1.0 + r1), 0.25 * (1.0 - r0) * (1.0 + r1)};
            const double d0[4] = {-0.25 * (1.0 - r1), 0.25 * (1.0 - r1), 0.25 * (1.0 + r1), -0.25 * (1.0 + r1)};
            const double d1[4] = {-0.25 * (1.0 - r0), -0.25 * (1.0 + r0), 0.25 * (1.0 + r0), 0.25 * (1.0 - r0)};
            double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
            for (int s = 0; s < 4; ++s)
              for (int d = 0; d < 3; ++d) {
                t1[d] += d0[s] * X[s][d];
                t2[d] += d1[s] * X[s][d];
              }
            const double cx = t1[1] * t2[2] - t1[2] * t2[1];
            const double cy = t1[2] * t2[0] - t1[0] * t2[2];
            const double cz = t1[0] * t2[1] - t1[1] * t2[0];
            const double jac = std::sqrt(cx * cx + cy * cy + cz * cz);
            for (int s = 0; s < 4; ++s)
              for (int c = 0; c < 3; ++c) fe[s][c] += N[s] * val[c] * jac;
          }
          for (int s = 0; s < 4; ++s)
            for (int c = 0; c < 3; ++c) {
              const int q = dof_[3 * nd[s] + c];
              if (q < numeq_) fext()[q] += fe[s][c];
            }
        }
      }
}

void Problem::filter_table(double radius, std::vector<int> &off, std::vector<double> &w,
                           int &R) const {
  if (!(radius > 0.0)) throw Error(kInvalidArgument, "filter radius must be positive");
  const double h[3] = {hx(), hy(), hz()};
  const int span = (int)std::ceil(radius / std::min(h[0], std::min(h[1], h[2])));
  off.clear();
  w.clear();
  R = 0;
  for (int dz = -span; dz <= span; ++dz)
    for (int dy = -span; dy <= span; ++dy)
      for (int dx = -span; dx <= span; ++dx) {
        const double a = dx * h[0], b = dy * h[1], c = dz * h[2];
        const double dist = std::sqrt(a * a + b * b + c * c);
        if (dist > radius) continue;
        off.push_back(dx);
        off.push_back(dy);
        off.push_back(dz);
        w.push_back(radius - dist);
        R = std::max(R, std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))));
      }
}

double gaussian_field(int nx, int ny, int nz, double h, uint64_t seed, int nblob, double smin,
                      double smax, double mean, double lo, double hi, double *out) {
  // splitmix64 stream, doubles from the top 53 bits
  uint64_t st = seed;
  auto next = [&]() {
    uint64_t z = (st += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  };
  auto uni = [&]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); };
  std::vector<double> amp(nblob), cx(nblob), cy(nblob), cz(nblob), s2(nblob);
  for (int b = 0; b < nblob; ++b) {
    amp[b] = 0.5 + uni();
    cx[b] = uni() * nx * h;
    cy[b] = uni() * ny * h;
    cz[b] = uni() * nz * h;
    const double s = smin + (smax - smin) * uni();
    s2[b] = 2.0 * s * s;
  }
  const long long n = (long long)nx * ny * nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double px = (i + 0.5) * h, py = (j + 0.5) * h, pz = (k + 0.5) * h;
        double v = 0.0;
        for (int b = 0; b < nblob; ++b) {
          const double dx = px - cx[b], dy = py - cy[b], dz = pz - cz[b];
          v += amp[b] * std::exp(-(dx * dx + dy * dy + dz * dz) / s2[b]);
        }
        out[i + (long long)nx * (j + (long long)ny * k)] = v;
      }
  double sum = 0.0;
  for (long long e = 0; e < n; ++e) sum += out[e];
  const double scale = mean / (sum / (double)n);
  double s = 0.0;
  for (long long e = 0; e < n; ++e) {
    double v = out[e] * scale;
    v = v < lo ? lo : (v > hi ? hi : v);
    out[e] = v;
    s += v;
  }
  return s / (double)n;
}

}  // namespace icarat_hip
