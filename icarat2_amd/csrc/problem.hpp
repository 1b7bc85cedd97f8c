// problem.hpp -- host-side description of one voxel hexa8 SIMP problem:
// lattice geometry, Dirichlet boxes and the reference DOF numbering, load
// vectors, the shared element stiffness KE0 and the filter stencil table.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "device_common.hpp"

namespace icarat_hip {

struct GridSpec {
  int nx, ny, nz;
  double lx, ly, lz;
  double young1, young2;  // [E1, E2] = [Emin, E0] (reference [material] young)
  double poisson, penalty;
};

class Problem {
 public:
  explicit Problem(const GridSpec &g);

  const GridSpec &spec() const { return g_; }
  const Lattice &lattice() const { return L_; }
  double hx() const { return g_.lx / g_.nx; }
  double hy() const { return g_.ly / g_.ny; }
  double hz() const { return g_.lz / g_.nz; }
  double coord(long long n, int d) const;

  // Dirichlet box with tolerance 1e-7 (reference
  // io/boundary_condition_icarat.hpp:72-138).
  void add_dirichlet(const int flag[6], const double val[6], const double range[6]);
  // the Dirichlet boxes as given (flag[6], val[6], range[6] each), for replaying
  // them on a coarser grid of the same box
  struct DirichletBox {
    int flag[6];
    double val[6], range[6];
  };
  const std::vector<DirichletBox> &dirichlet_boxes() const { return boxes_; }
  // Reference numbering: free DOFs first in node order, then the constrained
  // ones (reference io/boundary_condition_base.hpp:35-95).  Returns numeq.
  int make_dof();
  bool dof_ready() const { return numeq_ >= 0; }
  int numeq() const { return numeq_; }
  const std::vector<int32_t> &dof() const { return dof_; }
  const std::vector<uint8_t> &free_mask() const { return free_; }
  const std::vector<double> &dirichlet_values() const { return dirval_; }
  bool has_nonzero_dirichlet() const;

  // Point load (degenerate box) or face-distributed load (reference
  // io/boundary_condition_icarat.hpp:212-410).  Adds into the current load
  // case's fext (numeq order).
  void add_neumann(const double val[3], const double range[6]);
  std::vector<double> &fext() { return cases_[cur_]; }
  const std::vector<double> &fext() const { return cases_[cur_]; }
  // Load cases (one Force per case, reference opt_multi_load.cpp:74-84):
  // make_dof leaves one empty case; new_load_case appends one and makes it current.
  int new_load_case();
  void select_load_case(int i);
  int load_case() const { return cur_; }
  int num_load_cases() const { return (int)cases_.size(); }
  const std::vector<double> &case_fext(int i) const { return cases_[i]; }

  const double *ke0() const { return ke0_; }
  double element_volume() const { return vol_e_; }

  // Filter neighbourhood on the element lattice: offsets (dx,dy,dz) with
  // centre distance <= radius in ascending element-id order and weight
  // radius - distance (reference optimization/filter.hpp:254-297).
  void filter_table(double radius, std::vector<int> &off, std::vector<double> &w, int &R) const;

 private:
  GridSpec g_;
  Lattice L_;
  std::vector<DirichletBox> boxes_;
  std::vector<uint8_t> dirflag_;  // per node: bit c = constrained component c
  std::vector<uint8_t> hasdir_;
  std::vector<double> dirval_;    // numnp*3
  std::vector<int32_t> dof_;      // numnp*3
  std::vector<uint8_t> free_;     // per node: bit c = free component c
  std::vector<std::vector<double>> cases_{std::vector<double>()};
  int cur_ = 0;
  int numeq_ = -1;
  double ke0_[576];
  double vol_e_ = 0.0;
};

// Element stiffness of an axis-aligned hx*hy*hz hexa8 at young = 1 by the
// 2x2x2 Gauss rule (reference problem/elasticity.cpp:211-255,
// basis/bmatrix.hpp:90-162).  Returns the element volume.
double hexa8_stiffness(double hx, double hy, double hz, double nu, double ke[576]);

// Seeded smooth density field on the element lattice (sum of isotropic
// Gaussians rescaled to a target mean and clipped).  Returns the mean after
// clipping.
double gaussian_field(int nx, int ny, int nz, double h, uint64_t seed, int nblob, double smin,
                      double smax, double mean, double lo, double hi, double *out);

}  // namespace icarat_hip
