// host_problems.hpp -- host-side problem generators of the reference
// (blast.c, field_loop.c, cpaw2d.c) and the cpaw2d error norm, written on the
// reference's Grid geometry (init_grid.c:25-118, cc_pos.c:25-31) and layouts
// (Gas AoS of athena.h:43-52, interface fields [Nx2+8][Nx1+8]). Initial states
// use host libm with the reference's evaluation order, so they are identical
// to the reference's.
#pragma once
#include <cstddef>
#include <vector>

namespace athmi {

struct HostGeom {
  int Nx1, Nx2, is, ie, js, je, sx, sy, idisp, jdisp;
  double dx1, dx2, x1_0, x2_0, gamma_1;
  // cc_pos.c:25-31
  void cc_pos(int i, int j, double* x1, double* x2) const {
    *x1 = x1_0 + ((i + idisp) + 0.5) * dx1;
    *x2 = x2_0 + ((j + jdisp) + 0.5) * dx2;
  }
};

struct HostState {
  std::vector<double> gas, b1i, b2i, b3i;  // zero-initialised like calloc_2d_array
  explicit HostState(const HostGeom& G)
      : gas((size_t)8 * G.sx * G.sy, 0.0),
        b1i((size_t)G.sx * G.sy, 0.0),
        b2i((size_t)G.sx * G.sy, 0.0),
        b3i((size_t)G.sx * G.sy, 0.0) {}
};

enum GasVar { G_D = 0, G_M1, G_M2, G_M3, G_E, G_B1C, G_B2C, G_B3C };

void problem_blast(const HostGeom& G, HostState& S, double rin, double pa, double prat,
                   double b0, double angle_deg, double cx, double cy);
void problem_field_loop(const HostGeom& G, HostState& S, double rad, double amp, double vflow,
                        int iprob);
// cpaw2d.c:67-104: the wave's direction and wavenumber for a grid (host libm)
struct CpawConst {
  double sin_a, cos_a, b_par, b_perp, k_par, lambda;
};
CpawConst cpaw2d_constants(const HostGeom& G, double angle, double b_par, double b_perp);
double problem_cpaw2d(const HostGeom& G, HostState& S, double angle, double b_par, double b_perp,
                      double pres, double v_par, std::vector<double>* soln);
void cpaw2d_errors(const HostGeom& G, const HostState& S, const std::vector<double>& soln,
                   double* out);

}  // namespace athmi
