// host_problems.cpp -- see host_problems.hpp
#include "host_problems.hpp"

#include <cmath>

namespace athmi {

namespace {
constexpr double kPi = 3.14159265358979323846;  // defs.h:99
inline double sqr(double a) { return a * a; }
inline size_t at(const HostGeom& G, int j, int i) { return (size_t)j * G.sx + i; }
inline double& gas(HostState& S, const HostGeom& G, int j, int i, int q) {
  return S.gas[8 * at(G, j, i) + q];
}
}  // namespace

// blast.c:25-72. (cx, cy) displaces the over-pressured disc; the reference
// centres it at the origin (cx = cy = 0 reproduces it exactly).
void problem_blast(const HostGeom& G, HostState& S, double rin, double pa, double prat, double b0,
                   double angle_deg, double cx, double cy) {
  const double theta = (kPi / 180.0) * angle_deg;
  const double da = 1.0, ua = 0.0, va = 0.0, wa = 0.0;
  const double bxa = b0 * std::cos(theta);
  const double bya = b0 * std::sin(theta);
  const double bza = 0.0;
  for (int j = G.js; j <= G.je; j++) {
    for (int i = G.is; i <= G.ie; i++) {
      gas(S, G, j, i, G_D) = da;
      gas(S, G, j, i, G_M1) = da * ua;
      gas(S, G, j, i, G_M2) = da * va;
      gas(S, G, j, i, G_M3) = da * wa;
      S.b1i[at(G, j, i)] = bxa;
      S.b2i[at(G, j, i)] = bya;
      S.b3i[at(G, j, i)] = bza;
      gas(S, G, j, i, G_B1C) = bxa;
      gas(S, G, j, i, G_B2C) = bya;
      gas(S, G, j, i, G_B3C) = bza;
      if (i == G.ie && G.ie > G.is) S.b1i[at(G, j, i + 1)] = bxa;
      if (j == G.je && G.je > G.js) S.b2i[at(G, j + 1, i)] = bya;
      double x1, x2;
      G.cc_pos(i, j, &x1, &x2);
      const double rx = x1 - cx, ry = x2 - cy;
      const double rad = std::sqrt(rx * rx + ry * ry);
      double pressure = pa;
      if (rad < rin) pressure = prat * pa;
      gas(S, G, j, i, G_E) = pressure / G.gamma_1 + 0.5 * (bxa * bxa + bya * bya + bza * bza) +
                             0.5 * da * (ua * ua + va * va + wa * wa);
    }
  }
}

// field_loop.c:40-137 (iprob = 1: a loop of A_z = amp*(rad - r) in the x1-x2
// plane, advected with speed vflow along the box diagonal)
void problem_field_loop(const HostGeom& G, HostState& S, double rad, double amp, double vflow,
                        int iprob) {
  std::vector<double> az((size_t)G.sx * G.sy, 0.0);
  for (int j = G.js; j <= G.je + 1; j++) {
    for (int i = G.is; i <= G.ie + 1; i++) {
      double x1c, x2c;
      G.cc_pos(i, j, &x1c, &x2c);
      const double x1f = x1c - 0.5 * G.dx1;
      const double x2f = x2c - 0.5 * G.dx2;
      if (iprob == 1 && (x1f * x1f + x2f * x2f) < rad * rad)
        az[at(G, j, i)] = amp * (rad - std::sqrt(x1f * x1f + x2f * x2f));
    }
  }
  const double x1size = G.dx1 * (double)G.Nx1;
  const double x2size = G.dx2 * (double)G.Nx2;
  const double diag = std::sqrt(x1size * x1size + x2size * x2size);
  for (int j = G.js; j <= G.je; j++) {
    for (int i = G.is; i <= G.ie; i++) {
      const double d = 1.0;
      gas(S, G, j, i, G_D) = d;
      gas(S, G, j, i, G_M1) = d * vflow * x1size / diag;
      gas(S, G, j, i, G_M2) = d * vflow * x2size / diag;
      gas(S, G, j, i, G_M3) = 0.0;
      S.b1i[at(G, j, i)] = (az[at(G, j + 1, i)] - az[at(G, j, i)]) / G.dx2;
      S.b2i[at(G, j, i)] = -(az[at(G, j, i + 1)] - az[at(G, j, i)]) / G.dx1;
      S.b3i[at(G, j, i)] = 0.0;
    }
  }
  for (int j = G.js; j <= G.je; j++)
    S.b1i[at(G, j, G.ie + 1)] = (az[at(G, j + 1, G.ie + 1)] - az[at(G, j, G.ie + 1)]) / G.dx2;
  for (int i = G.is; i <= G.ie; i++)
    S.b2i[at(G, G.je + 1, i)] = -(az[at(G, G.je + 1, i + 1)] - az[at(G, G.je + 1, i)]) / G.dx1;
  for (int j = G.js; j <= G.je; j++) {
    for (int i = G.is; i <= G.ie; i++) {
      const double b1c = 0.5 * (S.b1i[at(G, j, i)] + S.b1i[at(G, j, i + 1)]);
      const double b2c = 0.5 * (S.b2i[at(G, j, i)] + S.b2i[at(G, j + 1, i)]);
      const double b3c = S.b3i[at(G, j, i)];
      gas(S, G, j, i, G_B1C) = b1c;
      gas(S, G, j, i, G_B2C) = b2c;
      gas(S, G, j, i, G_B3C) = b3c;
      const double m1 = gas(S, G, j, i, G_M1), m2 = gas(S, G, j, i, G_M2),
                   m3 = gas(S, G, j, i, G_M3), d = gas(S, G, j, i, G_D);
      gas(S, G, j, i, G_E) = 1.0 / G.gamma_1 + 0.5 * (sqr(b1c) + sqr(b2c) + sqr(b3c)) +
                             0.5 * (sqr(m1) + sqr(m2) + sqr(m3)) / d;
    }
  }
}

namespace {
// cpaw2d.c:311-315: B_x = -dA3/dy, B_y = dA3/dx
double A3(const CpawConst& c, double x1, double x2) {
  return c.b_par * (x1 * c.sin_a - x2 * c.cos_a) -
         (c.b_perp / c.k_par) * std::cos(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
}
}  // namespace

// one wavelength per direction, wave vector at atan(Lx/Ly) to x1 unless
// angle is 0 or 90
CpawConst cpaw2d_constants(const HostGeom& G, double angle, double b_par, double b_perp) {
  CpawConst c;
  double lambda;
  if (angle == 0.0) {
    c.sin_a = 0.0;
    c.cos_a = 1.0;
    lambda = G.Nx1 * G.dx1;
  } else if (angle == 90.0) {
    c.sin_a = 1.0;
    c.cos_a = 0.0;
    lambda = G.Nx2 * G.dx2;
  } else {
    if ((G.Nx1 * G.dx1) == (G.Nx2 * G.dx2)) {
      c.cos_a = c.sin_a = std::sqrt(0.5);
    } else {
      angle = std::atan((double)(G.Nx1 * G.dx1) / (G.Nx2 * G.dx2));
      c.sin_a = std::sin(angle);
      c.cos_a = std::cos(angle);
    }
    if (c.cos_a >= c.sin_a) lambda = G.Nx1 * G.dx1 * c.cos_a;
    else lambda = G.Nx2 * G.dx2 * c.sin_a;
  }
  c.lambda = lambda;
  c.k_par = 2.0 * kPi / lambda;
  c.b_par = b_par;
  c.b_perp = b_perp;
  return c;
}

// cpaw2d.c:56-185. soln receives the initial conserved state (8 planes
// [sy][sx], Gas order) for cpaw2d_errors.
double problem_cpaw2d(const HostGeom& G, HostState& S, double angle, double b_par, double b_perp,
                      double pres, double v_par, std::vector<double>* soln) {
  const CpawConst c = cpaw2d_constants(G, angle, b_par, b_perp);
  const double lambda = c.lambda;
  const double den = 1.0;
  const double v_perp = b_perp / std::sqrt(den);
  const double dx1 = G.dx1, dx2 = G.dx2;
  for (int j = G.js; j <= G.je + 1; j++) {
    for (int i = G.is; i <= G.ie + 1; i++) {
      double x1, x2;
      G.cc_pos(i, j, &x1, &x2);
      const double cs = std::cos(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
      x1 = x1 - 0.5 * G.dx1;
      x2 = x2 - 0.5 * G.dx2;
      S.b1i[at(G, j, i)] = -(A3(c, x1, x2 + dx2) - A3(c, x1, x2)) / dx2;
      S.b2i[at(G, j, i)] = (A3(c, x1 + dx1, x2) - A3(c, x1, x2)) / dx1;
      S.b3i[at(G, j, i)] = b_perp * cs;
    }
  }
  const size_t np = (size_t)G.sx * G.sy;
  if (soln) soln->assign(8 * np, 0.0);
  for (int j = G.js; j <= G.je; j++) {
    for (int i = G.is; i <= G.ie; i++) {
      double x1, x2;
      G.cc_pos(i, j, &x1, &x2);
      const double sn = std::sin(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
      const double cs = std::cos(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
      double v[8];
      v[G_D] = den;
      v[G_M1] = den * (v_par * c.cos_a - v_perp * sn * c.sin_a);
      v[G_M2] = den * (v_par * c.sin_a + v_perp * sn * c.cos_a);
      v[G_M3] = den * v_perp * cs;
      v[G_B1C] = 0.5 * (S.b1i[at(G, j, i)] + S.b1i[at(G, j, i + 1)]);
      v[G_B2C] = 0.5 * (S.b2i[at(G, j, i)] + S.b2i[at(G, j + 1, i)]);
      v[G_B3C] = b_perp * cs;
      v[G_E] = pres / G.gamma_1 + 0.5 * (sqr(v[G_B1C]) + sqr(v[G_B2C]) + sqr(v[G_B3C])) +
               0.5 * (sqr(v[G_M1]) + sqr(v[G_M2]) + sqr(v[G_M3])) / den;
      for (int q = 0; q < 8; q++) {
        gas(S, G, j, i, q) = v[q];
        if (soln) (*soln)[q * np + at(G, j, i)] = v[q];
      }
    }
  }
  return lambda;
}

// cpaw2d.c:222-302 (sums in j-major order; the RMS accumulates d, M1, M2, M3,
// then B1c, B2c, B3c, then E)
void cpaw2d_errors(const HostGeom& G, const HostState& S, const std::vector<double>& soln,
                   double* out) {
  const size_t np = (size_t)G.sx * G.sy;
  double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = G.js; j <= G.je; j++)
    for (int i = G.is; i <= G.ie; i++)
      for (int q = 0; q < 8; q++) {
        const size_t k = at(G, j, i);
        e[q] += std::fabs(S.gas[8 * k + q] - soln[q * np + k]);
      }
  double rms = sqr(e[G_D]) + sqr(e[G_M1]) + sqr(e[G_M2]) + sqr(e[G_M3]);
  rms += sqr(e[G_B1C]) + sqr(e[G_B2C]) + sqr(e[G_B3C]);
  rms += sqr(e[G_E]);
  const double n = (double)((G.ie - G.is + 1) * (G.je - G.js + 1));
  out[0] = std::sqrt(rms) / n;
  for (int q = 0; q < 8; q++) out[1 + q] = e[q] / n;
}

}  // namespace athmi
