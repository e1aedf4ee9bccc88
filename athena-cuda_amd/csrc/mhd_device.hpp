// mhd_device.hpp -- per-cell and per-face physics of the 2-D CTU+CT step for
// gfx950, written for registers: every 7x7 eigenmatrix is a set of named
// scalars with its structural zeros and sign-mirrored entries removed, so the
// projections cost one multiply per non-zero product and each product shared
// by two mirrored rows is formed once.
//
// Evaluation order follows the reference routines term by term (cited per
// routine); dropping a "+ 0*x" term or reusing an identical product does not
// change any finite result, so the kernels built with -ffp-contract=off agree
// with the reference CPU integrator value for value. The FMA-contracted
// variant (-ffp-contract=fast) differs only by the rounding of fused
// multiply-adds.
#pragma once
#include <hip/hip_runtime.h>

namespace athmi {

#define ATHMI_HD static __host__ __device__ __forceinline__
#define ATHMI_HDM __host__ __device__ __forceinline__  // member functions

constexpr double kTiny = 1.0e-20;  // defs.h:103 TINY_NUMBER

struct Cons7 { double d, mx, my, mz, e, by, bz; };  // athena.h:75-83 (Cons1D)
struct Prim7 { double d, vx, vy, vz, p, by, bz; };  // athena.h:89-97 (Prim1D)

// Physical constants of one grid (globals.h:18-31: Gamma, Gamma_1, Gamma_2)
struct EosC { double gamma, g1, g2; };

// defs.h:89-96 macro semantics
ATHMI_HD double dmax(double a, double b) { return (a > b) ? a : b; }
ATHMI_HD double dmin(double a, double b) { return (a < b) ? a : b; }
ATHMI_HD double dsign(double a) { return (a < 0.) ? -1. : 1.; }
ATHMI_HD double sqr(double a) { return a * a; }

// convert_var.c:37-55
ATHMI_HD Prim7 cons_to_prim(const Cons7& u, double bx, const EosC& eos) {
  Prim7 w;
  const double di = 1.0 / u.d;
  w.d = u.d;
  w.vx = u.mx * di;
  w.vy = u.my * di;
  w.vz = u.mz * di;
  double p = u.e - 0.5 * (sqr(u.mx) + sqr(u.my) + sqr(u.mz)) * di;
  p = p - 0.5 * (sqr(bx) + sqr(u.by) + sqr(u.bz));
  p = p * eos.g1;
  w.p = dmax(p, kTiny);
  w.by = u.by;
  w.bz = u.bz;
  return w;
}

// convert_var.c:64-92
ATHMI_HD Cons7 prim_to_cons(const Prim7& w, double bx, const EosC& eos) {
  Cons7 u;
  u.d = w.d;
  u.mx = w.d * w.vx;
  u.my = w.d * w.vy;
  u.mz = w.d * w.vz;
  double e = w.p / eos.g1 + 0.5 * w.d * (sqr(w.vx) + sqr(w.vy) + sqr(w.vz));
  u.e = e + 0.5 * (sqr(bx) + sqr(w.by) + sqr(w.bz));
  u.by = w.by;
  u.bz = w.bz;
  return u;
}

// ---------------------------------------------------------------------------
// Primitive-variable eigensystem (esystem_prim.c:36-235), non-zero entries only.
// Right eigenvectors are columns R[*][n]; left eigenvectors rows L[n][*].
struct EigPrim {
  double ev0, ev1, ev2, ev3, ev4, ev5, ev6;
  // R (mirrored entries: R04=R02, R06=R00, R14=-R12, R16=-R10, R24=-R22,
  // R26=-R20, R34=-R32, R36=-R30, R44=R42, R46=R40, R54=R52, R55=R51,
  // R56=R50, R64=R62, R65=R61, R66=R60; R03 = 1, R25 = bet3, R21 = -bet3,
  // R31 = bet2, R35 = -bet2)
  double R00, R02, R10, R12, R20, R22, R30, R32, R40, R42, R50, R51, R52, R60, R61, R62;
  double bet2, bet3;
  // L (L30 = 1; rows 4,5,6 mirror rows 2,1,0)
  double L01, L02, L03, L04, L05, L06, L12, L13, L15, L16;
  double L21, L22, L23, L24, L25, L26, L34;
};

ATHMI_HD void esys_prim(double d, double v1, double p, double b1, double b2, double b3,
                        const EosC& eos, EigPrim& E) {
  const double di = 1.0 / d;
  const double btsq = b2 * b2 + b3 * b3;
  const double vaxsq = b1 * b1 * di;
  const double asq = eos.gamma * p * di;
  const double ct2 = btsq * di;
  const double tsum = vaxsq + ct2 + asq;
  const double tdif = vaxsq + ct2 - asq;
  const double cf2_cs2 = sqrt(tdif * tdif + 4.0 * asq * ct2);
  const double cfsq = 0.5 * (tsum + cf2_cs2);
  const double cf = sqrt(cfsq);
  const double cssq = asq * vaxsq / cfsq;
  const double cs = sqrt(cssq);
  const double bt = sqrt(btsq);
  double bet2, bet3;
  if (bt == 0.0) {
    bet2 = 1.0;
    bet3 = 0.0;
  } else {
    bet2 = b2 / bt;
    bet3 = b3 / bt;
  }
  double alpha_f, alpha_s;
  if (cf2_cs2 == 0.0) {
    alpha_f = 1.0;
    alpha_s = 0.0;
  } else if ((asq - cssq) <= 0.0) {
    alpha_f = 0.0;
    alpha_s = 1.0;
  } else if ((cfsq - asq) <= 0.0) {
    alpha_f = 1.0;
    alpha_s = 0.0;
  } else {
    alpha_f = sqrt((asq - cssq) / cf2_cs2);
    alpha_s = sqrt((cfsq - asq) / cf2_cs2);
  }
  const double sqrtd = sqrt(d);
  const double s = dsign(b1);
  const double a = sqrt(asq);
  const double qf = cf * alpha_f * s;
  const double qs = cs * alpha_s * s;
  const double af = a * alpha_f * sqrtd;
  const double as = a * alpha_s * sqrtd;
  const double vax = sqrt(vaxsq);
  E.ev0 = v1 - cf;
  E.ev1 = v1 - vax;
  E.ev2 = v1 - cs;
  E.ev3 = v1;
  E.ev4 = v1 + cs;
  E.ev5 = v1 + vax;
  E.ev6 = v1 + cf;
  E.bet2 = bet2;
  E.bet3 = bet3;
  E.R00 = d * alpha_f;
  E.R02 = d * alpha_s;
  E.R10 = -cf * alpha_f;
  E.R12 = -cs * alpha_s;
  E.R20 = qs * bet2;
  E.R22 = -qf * bet2;
  E.R30 = qs * bet3;
  E.R32 = -qf * bet3;
  const double dasq = d * asq;
  E.R40 = dasq * alpha_f;
  E.R42 = dasq * alpha_s;
  E.R50 = as * bet2;
  E.R51 = -bet3 * s * sqrtd;
  E.R52 = -af * bet2;
  E.R60 = as * bet3;
  E.R61 = bet2 * s * sqrtd;
  E.R62 = -af * bet3;
  const double na = 0.5 / asq;
  const double qfn = na * qf;
  const double qsn = na * qs;
  const double af_prime = na * af * di;
  const double as_prime = na * as * di;
  E.L01 = -na * cf * alpha_f;
  E.L02 = qsn * bet2;
  E.L03 = qsn * bet3;
  E.L04 = na * alpha_f * di;
  E.L05 = as_prime * bet2;
  E.L06 = as_prime * bet3;
  E.L12 = -0.5 * bet3;
  E.L13 = 0.5 * bet2;
  E.L15 = -0.5 * bet3 * s / sqrtd;
  E.L16 = 0.5 * bet2 * s / sqrtd;
  E.L21 = -na * cs * alpha_s;
  E.L22 = -qfn * bet2;
  E.L23 = -qfn * bet3;
  E.L24 = na * alpha_s * di;
  E.L25 = -af_prime * bet2;
  E.L26 = -af_prime * bet3;
  E.L34 = -1.0 / asq;
}

// a = L * v (lr_states_plm.c:122-133 for one difference vector)
struct Vec7 { double c[7]; };

ATHMI_HD Vec7 project_L(const EigPrim& E, const double* v) {
  Vec7 a;
  // rows 0 and 6 share the products L0m*v[m] (L6m = -L0m for m=1..3, +L0m for 4..6)
  const double p1 = E.L01 * v[1], p2 = E.L02 * v[2], p3 = E.L03 * v[3];
  const double p4 = E.L04 * v[4], p5 = E.L05 * v[5], p6 = E.L06 * v[6];
  a.c[0] = p1 + p2 + p3 + p4 + p5 + p6;
  a.c[6] = -p1 - p2 - p3 + p4 + p5 + p6;
  // rows 1 and 5 (L5m = -L1m for m=2,3, +L1m for 5,6)
  const double q2 = E.L12 * v[2], q3 = E.L13 * v[3], q5 = E.L15 * v[5], q6 = E.L16 * v[6];
  a.c[1] = q2 + q3 + q5 + q6;
  a.c[5] = -q2 - q3 + q5 + q6;
  // rows 2 and 4
  const double r1 = E.L21 * v[1], r2 = E.L22 * v[2], r3 = E.L23 * v[3];
  const double r4 = E.L24 * v[4], r5 = E.L25 * v[5], r6 = E.L26 * v[6];
  a.c[2] = r1 + r2 + r3 + r4 + r5 + r6;
  a.c[4] = -r1 - r2 - r3 + r4 + r5 + r6;
  // row 3: L30 = 1
  a.c[3] = v[0] + E.L34 * v[4];
  return a;
}

// dWm = R * da (lr_states_plm.c:150-155)
ATHMI_HD Vec7 project_R(const EigPrim& E, const double* da) {
  Vec7 w;
  w.c[0] = E.R00 * da[0] + E.R02 * da[2] + da[3] + E.R02 * da[4] + E.R00 * da[6];
  w.c[1] = E.R10 * da[0] + E.R12 * da[2] + (-E.R12) * da[4] + (-E.R10) * da[6];
  w.c[2] = E.R20 * da[0] + (-E.bet3) * da[1] + E.R22 * da[2] + (-E.R22) * da[4] +
           E.bet3 * da[5] + (-E.R20) * da[6];
  w.c[3] = E.R30 * da[0] + E.bet2 * da[1] + E.R32 * da[2] + (-E.R32) * da[4] +
           (-E.bet2) * da[5] + (-E.R30) * da[6];
  w.c[4] = E.R40 * da[0] + E.R42 * da[2] + E.R42 * da[4] + E.R40 * da[6];
  w.c[5] = E.R50 * da[0] + E.R51 * da[1] + E.R52 * da[2] + E.R52 * da[4] + E.R51 * da[5] +
           E.R50 * da[6];
  w.c[6] = E.R60 * da[0] + E.R61 * da[1] + E.R62 * da[2] + E.R62 * da[4] + E.R61 * da[5] +
           E.R60 * da[6];
  return w;
}

// p[m] += q*R[m][N] (SGN = +1) or p[m] -= q*R[m][N] (SGN = -1) over the
// structurally non-zero entries of column N (lr_states_plm.c:207, 217, 222).
// A column entry that is the negation of a stored one uses the exact identity
// q*(-r) = -(q*r).
template <int SGN>
ATHMI_HD void acc(double& p, double t) { p = (SGN > 0) ? p + t : p - t; }
template <int SGN>
ATHMI_HD void acc_neg(double& p, double t) { p = (SGN > 0) ? p - t : p + t; }

template <int N, int SGN>
ATHMI_HD void axpy_col(const EigPrim& E, double q, double* p) {
  if constexpr (N == 0) {
    acc<SGN>(p[0], q * E.R00); acc<SGN>(p[1], q * E.R10); acc<SGN>(p[2], q * E.R20);
    acc<SGN>(p[3], q * E.R30); acc<SGN>(p[4], q * E.R40); acc<SGN>(p[5], q * E.R50);
    acc<SGN>(p[6], q * E.R60);
  } else if constexpr (N == 1) {
    acc_neg<SGN>(p[2], q * E.bet3); acc<SGN>(p[3], q * E.bet2);
    acc<SGN>(p[5], q * E.R51); acc<SGN>(p[6], q * E.R61);
  } else if constexpr (N == 2) {
    acc<SGN>(p[0], q * E.R02); acc<SGN>(p[1], q * E.R12); acc<SGN>(p[2], q * E.R22);
    acc<SGN>(p[3], q * E.R32); acc<SGN>(p[4], q * E.R42); acc<SGN>(p[5], q * E.R52);
    acc<SGN>(p[6], q * E.R62);
  } else if constexpr (N == 3) {
    acc<SGN>(p[0], q);
  } else if constexpr (N == 4) {
    acc<SGN>(p[0], q * E.R02); acc_neg<SGN>(p[1], q * E.R12); acc_neg<SGN>(p[2], q * E.R22);
    acc_neg<SGN>(p[3], q * E.R32); acc<SGN>(p[4], q * E.R42); acc<SGN>(p[5], q * E.R52);
    acc<SGN>(p[6], q * E.R62);
  } else if constexpr (N == 5) {
    acc<SGN>(p[2], q * E.bet3); acc_neg<SGN>(p[3], q * E.bet2);
    acc<SGN>(p[5], q * E.R51); acc<SGN>(p[6], q * E.R61);
  } else {
    acc<SGN>(p[0], q * E.R00); acc_neg<SGN>(p[1], q * E.R10); acc_neg<SGN>(p[2], q * E.R20);
    acc_neg<SGN>(p[3], q * E.R30); acc<SGN>(p[4], q * E.R40); acc<SGN>(p[5], q * E.R50);
    acc<SGN>(p[6], q * E.R60);
  }
}

// sum_m L[N][m]*(dtodx/2)*dev*dW[m] in ascending m, skipping structural zeros
// (lr_states_plm.c:204-205, 214-215, 219-220; lem*0.5*dtodx == lem*(0.5*dtodx)
// exactly since both scale by a power of two before the one rounding).
ATHMI_HD double tr(double l, double h, double dev, double w) { return l * h * dev * w; }

template <int N>
ATHMI_HD double trace_row(const EigPrim& E, double h, double dev, const double* dW) {
  if constexpr (N == 0) {
    return tr(E.L01, h, dev, dW[1]) + tr(E.L02, h, dev, dW[2]) + tr(E.L03, h, dev, dW[3]) +
           tr(E.L04, h, dev, dW[4]) + tr(E.L05, h, dev, dW[5]) + tr(E.L06, h, dev, dW[6]);
  } else if constexpr (N == 1) {
    return tr(E.L12, h, dev, dW[2]) + tr(E.L13, h, dev, dW[3]) + tr(E.L15, h, dev, dW[5]) +
           tr(E.L16, h, dev, dW[6]);
  } else if constexpr (N == 2) {
    return tr(E.L21, h, dev, dW[1]) + tr(E.L22, h, dev, dW[2]) + tr(E.L23, h, dev, dW[3]) +
           tr(E.L24, h, dev, dW[4]) + tr(E.L25, h, dev, dW[5]) + tr(E.L26, h, dev, dW[6]);
  } else if constexpr (N == 3) {
    return h * dev * dW[0] + tr(E.L34, h, dev, dW[4]);
  } else if constexpr (N == 4) {
    return tr(-E.L21, h, dev, dW[1]) + tr(-E.L22, h, dev, dW[2]) + tr(-E.L23, h, dev, dW[3]) +
           tr(E.L24, h, dev, dW[4]) + tr(E.L25, h, dev, dW[5]) + tr(E.L26, h, dev, dW[6]);
  } else if constexpr (N == 5) {
    return tr(-E.L12, h, dev, dW[2]) + tr(-E.L13, h, dev, dW[3]) + tr(E.L15, h, dev, dW[5]) +
           tr(E.L16, h, dev, dW[6]);
  } else {
    return tr(-E.L01, h, dev, dW[1]) + tr(-E.L02, h, dev, dW[2]) + tr(-E.L03, h, dev, dW[3]) +
           tr(E.L04, h, dev, dW[4]) + tr(E.L05, h, dev, dW[5]) + tr(E.L06, h, dev, dW[6]);
  }
}

template <int N>
ATHMI_HD void trace_right_moving(const EigPrim& E, double evn, double hdtodx, const double* dW,
                                 double* pWl) {
  if (evn > 0.) {
    const double qa = trace_row<N>(E, hdtodx, E.ev6 - evn, dW);
    axpy_col<N, +1>(E, qa, pWl);
  }
}

template <int N>
ATHMI_HD void trace_left_moving(const EigPrim& E, double evn, double hdtodx, const double* dW,
                                double* pWl, double* pWr) {
  if (evn < 0.) {
    const double qa = trace_row<N>(E, hdtodx, E.ev0 - evn, dW);
    axpy_col<N, +1>(E, qa, pWr);
    const double qb = trace_row<N>(E, hdtodx, evn - E.ev6, dW);
    axpy_col<N, -1>(E, qb, pWl);
  }
}

// ---------------------------------------------------------------------------
// lr_states_plm.c:95-245, loop body for one cell: from the primitive states
// of cells i-1, i, i+1 and the cell-centred normal field, the interface states
// Wl[i+1] (left of face i+1/2) and Wr[i] (right of face i-1/2).
// The stencil (cells i-1, i, i+1) is read through `S`: S.get(Wm, W0, Wp)
// fills the three states and S.fence() marks where earlier copies may die.
// RegStencil holds them in registers; a kernel that keeps the stencil in LDS
// passes a stencil that reads LDS after a fence, so the three states hold no
// registers while the eigen-projections run and are read again for the
// clamps. The arithmetic is the same either way.
struct RegStencil {
  const Prim7 &m, &z, &p;
  ATHMI_HDM void get(Prim7& a, Prim7& b, Prim7& c) const { a = m; b = z; c = p; }
  ATHMI_HDM void fence() const {}
};
template <class Stencil>
ATHMI_HD void plm_cell_s(const Stencil& S, double bxc, double dtodx, const EosC& eos,
                         Prim7& wl_next, Prim7& wr_this) {
  EigPrim E;
  double dWc[7], dWl[7], dWr[7], dWg[7];
  {
    Prim7 Wm, W0, Wp;
    S.get(Wm, W0, Wp);
    esys_prim(W0.d, W0.vx, W0.p, bxc, W0.by, W0.bz, eos, E);
    const double* wm = &Wm.d;
    const double* w0 = &W0.d;
    const double* wp = &Wp.d;
#pragma unroll
    for (int n = 0; n < 7; n++) {
      dWc[n] = wp[n] - wm[n];
      dWl[n] = w0[n] - wm[n];
      dWr[n] = wp[n] - w0[n];
      if (dWl[n] * dWr[n] > 0.0)
        dWg[n] = 2.0 * dWl[n] * dWr[n] / (dWl[n] + dWr[n]);
      else
        dWg[n] = 0.0;
    }
  }
  S.fence();
  const Vec7 dac = project_L(E, dWc);
  const Vec7 dal = project_L(E, dWl);
  const Vec7 dar = project_L(E, dWr);
  const Vec7 dag = project_L(E, dWg);
  double da[7];
#pragma unroll
  for (int n = 0; n < 7; n++) {
    da[n] = 0.0;
    if (dal.c[n] * dar.c[n] > 0.0) {
      const double lim1 = dmin(fabs(dal.c[n]), fabs(dar.c[n]));
      const double lim2 = dmin(0.5 * fabs(dac.c[n]), fabs(dag.c[n]));
      da[n] = dsign(dac.c[n]) * dmin(2.0 * lim1, lim2);
    }
  }
  const Vec7 dWm = project_R(E, da);
  S.fence();
  Prim7 Wm, W0, Wp;
  S.get(Wm, W0, Wp);
  const double* wm = &Wm.d;
  const double* w0 = &W0.d;
  const double* wp = &Wp.d;
  double Wlv[7], Wrv[7], dW[7];
#pragma unroll
  for (int n = 0; n < 7; n++) {
    double l = w0[n] - 0.5 * dWm.c[n];
    double r = w0[n] + 0.5 * dWm.c[n];
    l = dmax(dmin(w0[n], wm[n]), l);
    l = dmin(dmax(w0[n], wm[n]), l);
    r = dmax(dmin(w0[n], wp[n]), r);
    r = dmin(dmax(w0[n], wp[n]), r);
    Wlv[n] = l;
    Wrv[n] = r;
    dW[n] = r - l;
  }
  double pWl[7], pWr[7];
  double qx = 0.5 * dmax(E.ev6, 0.0) * dtodx;
#pragma unroll
  for (int n = 0; n < 7; n++) pWl[n] = Wrv[n] - qx * dW[n];
  qx = -0.5 * dmin(E.ev0, 0.0) * dtodx;
#pragma unroll
  for (int n = 0; n < 7; n++) pWr[n] = Wlv[n] + qx * dW[n];
  const double hdtodx = 0.5 * dtodx;
  // waves moving right do not all reach the interface (lr_states_plm.c:201-209)
  trace_right_moving<0>(E, E.ev0, hdtodx, dW, pWl);
  trace_right_moving<1>(E, E.ev1, hdtodx, dW, pWl);
  trace_right_moving<2>(E, E.ev2, hdtodx, dW, pWl);
  trace_right_moving<3>(E, E.ev3, hdtodx, dW, pWl);
  trace_right_moving<4>(E, E.ev4, hdtodx, dW, pWl);
  trace_right_moving<5>(E, E.ev5, hdtodx, dW, pWl);
  trace_right_moving<6>(E, E.ev6, hdtodx, dW, pWl);
  // waves moving left (lr_states_plm.c:211-224)
  trace_left_moving<0>(E, E.ev0, hdtodx, dW, pWl, pWr);
  trace_left_moving<1>(E, E.ev1, hdtodx, dW, pWl, pWr);
  trace_left_moving<2>(E, E.ev2, hdtodx, dW, pWl, pWr);
  trace_left_moving<3>(E, E.ev3, hdtodx, dW, pWl, pWr);
  trace_left_moving<4>(E, E.ev4, hdtodx, dW, pWl, pWr);
  trace_left_moving<5>(E, E.ev5, hdtodx, dW, pWl, pWr);
  trace_left_moving<6>(E, E.ev6, hdtodx, dW, pWl, pWr);
  wl_next = Prim7{pWl[0], pWl[1], pWl[2], pWl[3], pWl[4], pWl[5], pWl[6]};
  wr_this = Prim7{pWr[0], pWr[1], pWr[2], pWr[3], pWr[4], pWr[5], pWr[6]};
}
ATHMI_HD void plm_cell(const Prim7& Wm, const Prim7& W0, const Prim7& Wp, double bxc,
                       double dtodx, const EosC& eos, Prim7& wl_next, Prim7& wr_this) {
  plm_cell_s(RegStencil{Wm, W0, Wp}, bxc, dtodx, eos, wl_next, wr_this);
}

// ---------------------------------------------------------------------------
// Roe averages (flux_roe.c:91-114; identical in flux_hlle.c:87-110)
struct RoeAvg {
  double sqrtdl, sqrtdr, isdlpdr, droe, v1, v2, v3, b2, b3, x, y, pbl, pbr, h;
};

ATHMI_HD RoeAvg roe_average(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                            double Bxi) {
  RoeAvg r;
  r.sqrtdl = sqrt(Wl.d);
  r.sqrtdr = sqrt(Wr.d);
  r.isdlpdr = 1.0 / (r.sqrtdl + r.sqrtdr);
  r.droe = r.sqrtdl * r.sqrtdr;
  r.v1 = (r.sqrtdl * Wl.vx + r.sqrtdr * Wr.vx) * r.isdlpdr;
  r.v2 = (r.sqrtdl * Wl.vy + r.sqrtdr * Wr.vy) * r.isdlpdr;
  r.v3 = (r.sqrtdl * Wl.vz + r.sqrtdr * Wr.vz) * r.isdlpdr;
  r.b2 = (r.sqrtdr * Wl.by + r.sqrtdl * Wr.by) * r.isdlpdr;
  r.b3 = (r.sqrtdr * Wl.bz + r.sqrtdl * Wr.bz) * r.isdlpdr;
  r.x = 0.5 * (sqr(Wl.by - Wr.by) + sqr(Wl.bz - Wr.bz)) / (sqr(r.sqrtdl + r.sqrtdr));
  r.y = 0.5 * (Wl.d + Wr.d) / r.droe;
  r.pbl = 0.5 * (sqr(Bxi) + sqr(Wl.by) + sqr(Wl.bz));
  r.pbr = 0.5 * (sqr(Bxi) + sqr(Wr.by) + sqr(Wr.bz));
  r.h = ((Ul.e + Wl.p + r.pbl) / r.sqrtdl + (Ur.e + Wr.p + r.pbr) / r.sqrtdr) * r.isdlpdr;
  return r;
}

// Shared scalars of the Roe-matrix eigensystem (esystem_roe.c:68-145)
struct RoeSpeeds {
  double ev0, ev1, ev2, ev3, ev4, ev5, ev6;
  double vsq, hp, twid_asq, cf, cs, bt_star, bet2, bet3, bet2_star, bet3_star, bet_starsq,
      vbet, alpha_f, alpha_s, sqrtd, isqrtd, s, qf, qs, af_prime, as_prime;
};

ATHMI_HD void roe_speeds(double d, double v1, double v2, double v3, double h, double b1,
                         double b2, double b3, double x, double y, const EosC& eos,
                         bool need_vectors, RoeSpeeds& S) {
  const double Gamma_1 = eos.g1, Gamma_2 = eos.g2;
  const double di = 1.0 / d;
  const double vsq = v1 * v1 + v2 * v2 + v3 * v3;
  const double btsq = b2 * b2 + b3 * b3;
  const double bt_starsq = (Gamma_1 - Gamma_2 * y) * btsq;
  const double vaxsq = b1 * b1 * di;
  const double hp = h - (vaxsq + btsq * di);
  const double twid_asq = dmax((Gamma_1 * (hp - 0.5 * vsq) - Gamma_2 * x), kTiny);
  const double ct2 = bt_starsq * di;
  const double tsum = vaxsq + ct2 + twid_asq;
  const double tdif = vaxsq + ct2 - twid_asq;
  const double cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
  const double cfsq = 0.5 * (tsum + cf2_cs2);
  const double cf = sqrt(cfsq);
  const double cssq = twid_asq * vaxsq / cfsq;
  const double cs = sqrt(cssq);
  const double vax = sqrt(vaxsq);
  S.ev0 = v1 - cf;
  S.ev1 = v1 - vax;
  S.ev2 = v1 - cs;
  S.ev3 = v1;
  S.ev4 = v1 + cs;
  S.ev5 = v1 + vax;
  S.ev6 = v1 + cf;
  S.cf = cf;
  S.cs = cs;
  if (!need_vectors) return;
  S.vsq = vsq;
  S.hp = hp;
  S.twid_asq = twid_asq;
  const double bt = sqrt(btsq);
  S.bt_star = sqrt(bt_starsq);
  double bet2, bet3;
  if (bt == 0.0) {
    bet2 = 1.0;
    bet3 = 0.0;
  } else {
    bet2 = b2 / bt;
    bet3 = b3 / bt;
  }
  S.bet2 = bet2;
  S.bet3 = bet3;
  const double sq = sqrt(Gamma_1 - Gamma_2 * y);
  S.bet2_star = bet2 / sq;
  S.bet3_star = bet3 / sq;
  S.bet_starsq = S.bet2_star * S.bet2_star + S.bet3_star * S.bet3_star;
  S.vbet = v2 * S.bet2_star + v3 * S.bet3_star;
  double alpha_f, alpha_s;
  if ((cfsq - cssq) == 0.0) {
    alpha_f = 1.0;
    alpha_s = 0.0;
  } else if ((twid_asq - cssq) <= 0.0) {
    alpha_f = 0.0;
    alpha_s = 1.0;
  } else if ((cfsq - twid_asq) <= 0.0) {
    alpha_f = 1.0;
    alpha_s = 0.0;
  } else {
    alpha_f = sqrt((twid_asq - cssq) / (cfsq - cssq));
    alpha_s = sqrt((cfsq - twid_asq) / (cfsq - cssq));
  }
  S.alpha_f = alpha_f;
  S.alpha_s = alpha_s;
  S.sqrtd = sqrt(d);
  S.isqrtd = 1.0 / S.sqrtd;
  S.s = dsign(b1);
  const double twid_a = sqrt(twid_asq);
  S.qf = cf * alpha_f * S.s;
  S.qs = cs * alpha_s * S.s;
  S.af_prime = twid_a * alpha_f * S.isqrtd;
  S.as_prime = twid_a * alpha_s * S.isqrtd;
}

// Left eigenvectors of the Roe matrix (esystem_roe.c:218-291), non-zero part
struct RoeL {
  double L00, L01, L02, L03, L04, L05, L06;
  double L10, L12, L13, L15, L16;  // row 5 mirrors row 1
  double L20, L21, L22, L23, L24, L25, L26;
  double L30, L31, L32, L33, L34, L35, L36;
  double L40, L41, L42, L43, L44;  // L45 = L25, L46 = L26
  double L60, L61, L62, L63, L64;  // L65 = L05, L66 = L06
};

ATHMI_HD void roe_left(const RoeSpeeds& S, double d, double v1, double v2, double v3, double b2,
                       double b3, double x, const EosC& eos, RoeL& L) {
  const double Gamma_1 = eos.g1, Gamma_2 = eos.g2;
  double norm = 0.5 / S.twid_asq;
  const double cff = norm * S.alpha_f * S.cf;
  const double css = norm * S.alpha_s * S.cs;
  const double qf = S.qf * norm;
  const double qs = S.qs * norm;
  const double af = norm * S.af_prime * d;
  const double as = norm * S.as_prime * d;
  const double afpb = norm * S.af_prime * S.bt_star;
  const double aspb = norm * S.as_prime * S.bt_star;
  norm = norm * Gamma_1;
  const double alpha_f = S.alpha_f * norm;
  const double alpha_s = S.alpha_s * norm;
  const double q2_star = S.bet2_star / S.bet_starsq;
  const double q3_star = S.bet3_star / S.bet_starsq;
  const double vqstr = (v2 * q2_star + v3 * q3_star);
  norm = norm * 2.0;
  const double vsq = S.vsq, hp = S.hp, cf = S.cf, cs = S.cs;
  L.L00 = alpha_f * (vsq - hp) + cff * (cf + v1) - qs * vqstr - aspb;
  L.L01 = -alpha_f * v1 - cff;
  L.L02 = -alpha_f * v2 + qs * q2_star;
  L.L03 = -alpha_f * v3 + qs * q3_star;
  L.L04 = alpha_f;
  L.L05 = as * q2_star - alpha_f * b2;
  L.L06 = as * q3_star - alpha_f * b3;
  L.L10 = 0.5 * (v2 * S.bet3 - v3 * S.bet2);
  L.L12 = -0.5 * S.bet3;
  L.L13 = 0.5 * S.bet2;
  L.L15 = -0.5 * S.sqrtd * S.bet3 * S.s;
  L.L16 = 0.5 * S.sqrtd * S.bet2 * S.s;
  L.L20 = alpha_s * (vsq - hp) + css * (cs + v1) + qf * vqstr + afpb;
  L.L21 = -alpha_s * v1 - css;
  L.L22 = -alpha_s * v2 - qf * q2_star;
  L.L23 = -alpha_s * v3 - qf * q3_star;
  L.L24 = alpha_s;
  L.L25 = -af * q2_star - alpha_s * b2;
  L.L26 = -af * q3_star - alpha_s * b3;
  L.L30 = 1.0 - norm * (0.5 * vsq - Gamma_2 * x / Gamma_1);
  L.L31 = norm * v1;
  L.L32 = norm * v2;
  L.L33 = norm * v3;
  L.L34 = -norm;
  L.L35 = norm * b2;
  L.L36 = norm * b3;
  L.L40 = alpha_s * (vsq - hp) + css * (cs - v1) - qf * vqstr + afpb;
  L.L41 = -alpha_s * v1 + css;
  L.L42 = -alpha_s * v2 + qf * q2_star;
  L.L43 = -alpha_s * v3 + qf * q3_star;
  L.L44 = alpha_s;
  L.L60 = alpha_f * (vsq - hp) + cff * (cf - v1) + qs * vqstr - aspb;
  L.L61 = -alpha_f * v1 + cff;
  L.L62 = -alpha_f * v2 - qs * q2_star;
  L.L63 = -alpha_f * v3 - qs * q3_star;
  L.L64 = alpha_f;
}

// Right eigenvectors of the Roe matrix (esystem_roe.c:151-213), non-zero part
struct RoeR {
  double R00, R02;  // R04 = R02, R06 = R00, R03 = 1
  double R10, R12, R13, R14, R16;
  double R20, R21, R22, R23, R24, R25, R26;
  double R30, R31, R32, R33, R34, R35, R36;
  double R40, R41, R42, R43, R44, R46;  // R45 = -R41
  double R50, R51, R52;  // R54 = R52, R55 = R51, R56 = R50
  double R60, R61, R62;  // R64 = R62, R65 = R61, R66 = R60
};

ATHMI_HD void roe_right(const RoeSpeeds& S, double v1, double v2, double v3, double x,
                        const EosC& eos, RoeR& R) {
  const double Gamma_1 = eos.g1, Gamma_2 = eos.g2;
  const double alpha_f = S.alpha_f, alpha_s = S.alpha_s;
  const double afpbb = S.af_prime * S.bt_star * S.bet_starsq;
  const double aspbb = S.as_prime * S.bt_star * S.bet_starsq;
  R.R00 = alpha_f;
  R.R02 = alpha_s;
  R.R10 = alpha_f * S.ev0;
  R.R12 = alpha_s * S.ev2;
  R.R13 = v1;
  R.R14 = alpha_s * S.ev4;
  R.R16 = alpha_f * S.ev6;
  double qa = alpha_f * v2, qb = alpha_s * v2, qc = S.qs * S.bet2_star, qd = S.qf * S.bet2_star;
  R.R20 = qa + qc;
  R.R21 = -S.bet3;
  R.R22 = qb - qd;
  R.R23 = v2;
  R.R24 = qb + qd;
  R.R25 = S.bet3;
  R.R26 = qa - qc;
  qa = alpha_f * v3;
  qb = alpha_s * v3;
  qc = S.qs * S.bet3_star;
  qd = S.qf * S.bet3_star;
  R.R30 = qa + qc;
  R.R31 = S.bet2;
  R.R32 = qb - qd;
  R.R33 = v3;
  R.R34 = qb + qd;
  R.R35 = -S.bet2;
  R.R36 = qa - qc;
  const double hp = S.hp;
  R.R40 = alpha_f * (hp - v1 * S.cf) + S.qs * S.vbet + aspbb;
  R.R41 = -(v2 * S.bet3 - v3 * S.bet2);
  R.R42 = alpha_s * (hp - v1 * S.cs) - S.qf * S.vbet - afpbb;
  R.R43 = 0.5 * S.vsq + Gamma_2 * x / Gamma_1;
  R.R44 = alpha_s * (hp + v1 * S.cs) + S.qf * S.vbet - afpbb;
  R.R46 = alpha_f * (hp + v1 * S.cf) - S.qs * S.vbet + aspbb;
  R.R50 = S.as_prime * S.bet2_star;
  R.R51 = -S.bet3 * S.s * S.isqrtd;
  R.R52 = -S.af_prime * S.bet2_star;
  R.R60 = S.as_prime * S.bet3_star;
  R.R61 = S.bet2 * S.s * S.isqrtd;
  R.R62 = -S.af_prime * S.bet3_star;
}

// u[m] += q*R[m][N] for the Roe-matrix column N (flux_roe.c:217)
template <int N>
ATHMI_HD void axpy_roe_col(const RoeR& R, double q, double* u) {
  if constexpr (N == 0) {
    u[0] += q * R.R00; u[1] += q * R.R10; u[2] += q * R.R20; u[3] += q * R.R30;
    u[4] += q * R.R40; u[5] += q * R.R50; u[6] += q * R.R60;
  } else if constexpr (N == 1) {
    u[2] += q * R.R21; u[3] += q * R.R31; u[4] += q * R.R41; u[5] += q * R.R51; u[6] += q * R.R61;
  } else if constexpr (N == 2) {
    u[0] += q * R.R02; u[1] += q * R.R12; u[2] += q * R.R22; u[3] += q * R.R32;
    u[4] += q * R.R42; u[5] += q * R.R52; u[6] += q * R.R62;
  } else if constexpr (N == 3) {
    u[0] += q; u[1] += q * R.R13; u[2] += q * R.R23; u[3] += q * R.R33; u[4] += q * R.R43;
  } else if constexpr (N == 4) {
    u[0] += q * R.R02; u[1] += q * R.R14; u[2] += q * R.R24; u[3] += q * R.R34;
    u[4] += q * R.R44; u[5] += q * R.R52; u[6] += q * R.R62;
  } else {  // N == 5
    u[2] += q * R.R25; u[3] += q * R.R35; u[4] -= q * R.R41; u[5] += q * R.R51; u[6] += q * R.R61;
  }
}

// one pass of the intermediate-state test (flux_roe.c:216-231); returns the
// hlle_flag it raises (0 while the states stay physical)
template <int N>
ATHMI_HD int inter_state(const RoeR& R, const double* a, const double* ev, double Bxi, double* ui) {
  axpy_roe_col<N>(R, a[N], ui);
  if (ev[N + 1] > ev[N]) {
    if (ui[0] <= 0.0) return 1;
    double p_inter = ui[4] - 0.5 * (sqr(ui[1]) + sqr(ui[2]) + sqr(ui[3])) / ui[0];
    p_inter = p_inter - 0.5 * (sqr(ui[5]) + sqr(ui[6]) + sqr(Bxi));
    if (p_inter < 0.0) return 2;
  }
  return 0;
}

// Physical fluxes of the L/R states (flux_roe.c:136-172)
ATHMI_HD Cons7 side_flux(const Cons7& U, const Prim7& W, double Bxi, double pb) {
  Cons7 F;
  F.d = U.mx;
  F.mx = U.mx * W.vx;
  F.my = U.mx * W.vy;
  F.mz = U.mx * W.vz;
  F.mx = F.mx + W.p;
  F.e = (U.e + W.p) * W.vx;
  F.mx = F.mx - 0.5 * (Bxi * Bxi - sqr(W.by) - sqr(W.bz));
  F.my = F.my - Bxi * W.by;
  F.mz = F.mz - Bxi * W.bz;
  F.e = F.e + (pb * W.vx - Bxi * (Bxi * W.vx + W.by * W.vy + W.bz * W.vz));
  F.by = W.by * W.vx - Bxi * W.vy;
  F.bz = W.bz * W.vx - Bxi * W.vz;
  return F;
}

// flux_hlle.c:118-196 given the Roe averages and the Roe eigenvalues
ATHMI_HD Cons7 hlle_from_roe(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                             double Bxi, const RoeAvg& r, double ev0, double ev6,
                             const EosC& eos) {
  double asq = eos.gamma * Wl.p / Wl.d;
  double vaxsq = Bxi * Bxi / Wl.d;
  double ct2 = (Ul.by * Ul.by + Ul.bz * Ul.bz) / Wl.d;
  double qsq = vaxsq + ct2 + asq;
  double tmp = vaxsq + ct2 - asq;
  double cfsq = 0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2));
  const double cfl = sqrt(cfsq);
  asq = eos.gamma * Wr.p / Wr.d;
  vaxsq = Bxi * Bxi / Wr.d;
  ct2 = (Ur.by * Ur.by + Ur.bz * Ur.bz) / Wr.d;
  qsq = vaxsq + ct2 + asq;
  tmp = vaxsq + ct2 - asq;
  cfsq = 0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2));
  const double cfr = sqrt(cfsq);
  const double ar = dmax(ev6, (Wr.vx + cfr));
  const double al = dmin(ev0, (Wl.vx - cfl));
  const double bp = dmax(ar, 0.0);
  const double bm = dmin(al, 0.0);
  Cons7 Fl, Fr;
  Fl.d = Ul.mx - bm * Ul.d;
  Fr.d = Ur.mx - bp * Ur.d;
  Fl.mx = Ul.mx * (Wl.vx - bm);
  Fr.mx = Ur.mx * (Wr.vx - bp);
  Fl.my = Ul.my * (Wl.vx - bm);
  Fr.my = Ur.my * (Wr.vx - bp);
  Fl.mz = Ul.mz * (Wl.vx - bm);
  Fr.mz = Ur.mz * (Wr.vx - bp);
  Fl.mx = Fl.mx + Wl.p;
  Fr.mx = Fr.mx + Wr.p;
  Fl.e = Ul.e * (Wl.vx - bm) + Wl.p * Wl.vx;
  Fr.e = Ur.e * (Wr.vx - bp) + Wr.p * Wr.vx;
  Fl.mx = Fl.mx - 0.5 * (Bxi * Bxi - sqr(Wl.by) - sqr(Wl.bz));
  Fr.mx = Fr.mx - 0.5 * (Bxi * Bxi - sqr(Wr.by) - sqr(Wr.bz));
  Fl.my = Fl.my - Bxi * Wl.by;
  Fr.my = Fr.my - Bxi * Wr.by;
  Fl.mz = Fl.mz - Bxi * Wl.bz;
  Fr.mz = Fr.mz - Bxi * Wr.bz;
  Fl.e = Fl.e + (r.pbl * Wl.vx - Bxi * (Bxi * Wl.vx + Wl.by * Wl.vy + Wl.bz * Wl.vz));
  Fr.e = Fr.e + (r.pbr * Wr.vx - Bxi * (Bxi * Wr.vx + Wr.by * Wr.vy + Wr.bz * Wr.vz));
  Fl.by = Wl.by * (Wl.vx - bm) - Bxi * Wl.vy;
  Fr.by = Wr.by * (Wr.vx - bp) - Bxi * Wr.vy;
  Fl.bz = Wl.bz * (Wl.vx - bm) - Bxi * Wl.vz;
  Fr.bz = Wr.bz * (Wr.vx - bp) - Bxi * Wr.vz;
  tmp = 0.5 * (bp + bm) / (bp - bm);
  const double* fl = &Fl.d;
  const double* fr = &Fr.d;
  Cons7 F;
  double* f = &F.d;
#pragma unroll
  for (int n = 0; n < 7; n++) f[n] = 0.5 * (fl[n] + fr[n]) + (fl[n] - fr[n]) * tmp;
  return F;
}

// flux_hlle.c:60-199
ATHMI_HD Cons7 flux_hlle(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                         double Bxi, const EosC& eos) {
  const RoeAvg r = roe_average(Ul, Ur, Wl, Wr, Bxi);
  RoeSpeeds S;
  roe_speeds(r.droe, r.v1, r.v2, r.v3, r.h, Bxi, r.b2, r.b3, r.x, r.y, eos, false, S);
  return hlle_from_roe(Ul, Ur, Wl, Wr, Bxi, r, S.ev0, S.ev6, eos);
}

// Scalars of the Roe-matrix eigenvectors that are shared by several rows
// or columns (esystem_roe.c:129-132 and :218-235)
struct RoeVecScalars {
  double afpbb, aspbb;                             // right-eigenvector terms
  double cff, css, qfn, qsn, afn, asn, afpb, aspb; // left, normalised by 1/2a^2
  double alpha_fn, alpha_sn, q2_star, q3_star, vqstr, norm2;
};

ATHMI_HD void roe_vec_scalars(const RoeSpeeds& S, double d, double v2, double v3,
                              const EosC& eos, RoeVecScalars& V) {
  V.afpbb = S.af_prime * S.bt_star * S.bet_starsq;
  V.aspbb = S.as_prime * S.bt_star * S.bet_starsq;
  double norm = 0.5 / S.twid_asq;
  V.cff = norm * S.alpha_f * S.cf;
  V.css = norm * S.alpha_s * S.cs;
  V.qfn = S.qf * norm;
  V.qsn = S.qs * norm;
  V.afn = norm * S.af_prime * d;
  V.asn = norm * S.as_prime * d;
  V.afpb = norm * S.af_prime * S.bt_star;
  V.aspb = norm * S.as_prime * S.bt_star;
  norm = norm * eos.g1;
  V.alpha_fn = S.alpha_f * norm;
  V.alpha_sn = S.alpha_s * norm;
  V.q2_star = S.bet2_star / S.bet_starsq;
  V.q3_star = S.bet3_star / S.bet_starsq;
  V.vqstr = (v2 * V.q2_star + v3 * V.q3_star);
  V.norm2 = norm * 2.0;
}

// a = L dU with each left-eigenvector row formed where it is used
// (esystem_roe.c:237-291; flux_roe.c:200-204 sums m = 0..6 in order)
ATHMI_HD void roe_project(const RoeSpeeds& S, const RoeVecScalars& V, double v1, double v2,
                          double v3, double b2, double b3, double x, const EosC& eos,
                          const double* du, double* a) {
  const double vsq = S.vsq, hp = S.hp, cf = S.cf, cs = S.cs;
  const double af = V.alpha_fn, as_ = V.alpha_sn;
  // rows 0 and 6 share L05 = L65, L06 = L66
  const double L05 = V.asn * V.q2_star - af * b2;
  const double L06 = V.asn * V.q3_star - af * b3;
  const double p5 = L05 * du[5], p6 = L06 * du[6];
  a[0] = (af * (vsq - hp) + V.cff * (cf + v1) - V.qsn * V.vqstr - V.aspb) * du[0] +
         (-af * v1 - V.cff) * du[1] + (-af * v2 + V.qsn * V.q2_star) * du[2] +
         (-af * v3 + V.qsn * V.q3_star) * du[3] + af * du[4] + p5 + p6;
  a[6] = (af * (vsq - hp) + V.cff * (cf - v1) + V.qsn * V.vqstr - V.aspb) * du[0] +
         (-af * v1 + V.cff) * du[1] + (-af * v2 - V.qsn * V.q2_star) * du[2] +
         (-af * v3 - V.qsn * V.q3_star) * du[3] + af * du[4] + p5 + p6;
  // rows 1 and 5
  const double q0 = (0.5 * (v2 * S.bet3 - v3 * S.bet2)) * du[0];
  const double q2 = (-0.5 * S.bet3) * du[2], q3 = (0.5 * S.bet2) * du[3];
  const double q5 = (-0.5 * S.sqrtd * S.bet3 * S.s) * du[5];
  const double q6 = (0.5 * S.sqrtd * S.bet2 * S.s) * du[6];
  a[1] = q0 + q2 + q3 + q5 + q6;
  a[5] = -q0 - q2 - q3 + q5 + q6;
  // rows 2 and 4 share L25 = L45, L26 = L46
  const double s5 = (-V.afn * V.q2_star - as_ * b2) * du[5];
  const double s6 = (-V.afn * V.q3_star - as_ * b3) * du[6];
  a[2] = (as_ * (vsq - hp) + V.css * (cs + v1) + V.qfn * V.vqstr + V.afpb) * du[0] +
         (-as_ * v1 - V.css) * du[1] + (-as_ * v2 - V.qfn * V.q2_star) * du[2] +
         (-as_ * v3 - V.qfn * V.q3_star) * du[3] + as_ * du[4] + s5 + s6;
  a[4] = (as_ * (vsq - hp) + V.css * (cs - v1) - V.qfn * V.vqstr + V.afpb) * du[0] +
         (-as_ * v1 + V.css) * du[1] + (-as_ * v2 + V.qfn * V.q2_star) * du[2] +
         (-as_ * v3 + V.qfn * V.q3_star) * du[3] + as_ * du[4] + s5 + s6;
  // row 3
  const double n2 = V.norm2;
  a[3] = (1.0 - n2 * (0.5 * vsq - eos.g2 * x / eos.g1)) * du[0] + (n2 * v1) * du[1] +
         (n2 * v2) * du[2] + (n2 * v3) * du[3] + (-n2) * du[4] + (n2 * b2) * du[5] +
         (n2 * b3) * du[6];
}

// column N of the Roe-matrix R (esystem_roe.c:151-213); r[m] = R[m][N]
template <int N>
ATHMI_HD void roe_col(const RoeSpeeds& S, const RoeVecScalars& V, double v1, double v2,
                      double v3, double x, const EosC& eos, double* r) {
  const double af = S.alpha_f, as_ = S.alpha_s, hp = S.hp;
  if constexpr (N == 0) {
    r[0] = af; r[1] = af * S.ev0;
    r[2] = af * v2 + S.qs * S.bet2_star; r[3] = af * v3 + S.qs * S.bet3_star;
    r[4] = af * (hp - v1 * S.cf) + S.qs * S.vbet + V.aspbb;
    r[5] = S.as_prime * S.bet2_star; r[6] = S.as_prime * S.bet3_star;
  } else if constexpr (N == 1) {
    r[0] = 0.0; r[1] = 0.0; r[2] = -S.bet3; r[3] = S.bet2;
    r[4] = -(v2 * S.bet3 - v3 * S.bet2);
    r[5] = -S.bet3 * S.s * S.isqrtd; r[6] = S.bet2 * S.s * S.isqrtd;
  } else if constexpr (N == 2) {
    r[0] = as_; r[1] = as_ * S.ev2;
    r[2] = as_ * v2 - S.qf * S.bet2_star; r[3] = as_ * v3 - S.qf * S.bet3_star;
    r[4] = as_ * (hp - v1 * S.cs) - S.qf * S.vbet - V.afpbb;
    r[5] = -S.af_prime * S.bet2_star; r[6] = -S.af_prime * S.bet3_star;
  } else if constexpr (N == 3) {
    r[0] = 1.0; r[1] = v1; r[2] = v2; r[3] = v3;
    r[4] = 0.5 * S.vsq + eos.g2 * x / eos.g1; r[5] = 0.0; r[6] = 0.0;
  } else if constexpr (N == 4) {
    r[0] = as_; r[1] = as_ * S.ev4;
    r[2] = as_ * v2 + S.qf * S.bet2_star; r[3] = as_ * v3 + S.qf * S.bet3_star;
    r[4] = as_ * (hp + v1 * S.cs) + S.qf * S.vbet - V.afpbb;
    r[5] = -S.af_prime * S.bet2_star; r[6] = -S.af_prime * S.bet3_star;
  } else if constexpr (N == 5) {
    r[0] = 0.0; r[1] = 0.0; r[2] = S.bet3; r[3] = -S.bet2;
    r[4] = (v2 * S.bet3 - v3 * S.bet2);
    r[5] = -S.bet3 * S.s * S.isqrtd; r[6] = S.bet2 * S.s * S.isqrtd;
  } else {
    r[0] = af; r[1] = af * S.ev6;
    r[2] = af * v2 - S.qs * S.bet2_star; r[3] = af * v3 - S.qs * S.bet3_star;
    r[4] = af * (hp + v1 * S.cf) - S.qs * S.vbet + V.aspbb;
    r[5] = S.as_prime * S.bet2_star; r[6] = S.as_prime * S.bet3_star;
  }
}

// structural non-zeros of the columns (R01, R05, R11, R15, R53, R63 are zero)
template <int N>
ATHMI_HD constexpr bool roe_nz(int m) {
  return !((N == 1 || N == 5) && (m == 0 || m == 1)) && !(N == 3 && (m == 5 || m == 6));
}

// One column of the Roe-flux sum: the intermediate state u += a[N]*R[*][N] and
// its positivity test (flux_roe.c:216-231), and f -= coeff[N]*R[*][N]
// (flux_roe.c:251-256; per component the sum still runs over N ascending).
// The test is evaluated without branches: the reference stops at the first
// failing column and takes the HLLE flux, so only whether any column fails
// matters (a state after a failure is never used).
template <int N>
ATHMI_HD void roe_column_pass(const RoeSpeeds& S, const RoeVecScalars& V, double v1,
                              double v2, double v3, double x, const EosC& eos, const double* a,
                              const double* coeff, const double* ev, double Bxi, double* ui,
                              double* f, bool& bad) {
  double r[7];
  roe_col<N>(S, V, v1, v2, v3, x, eos, r);
  if constexpr (N < 6) {
#pragma unroll
    for (int m = 0; m < 7; m++)
      if (roe_nz<N>(m)) ui[m] = ui[m] + (N == 3 && m == 0 ? a[N] : a[N] * r[m]);
    double p_inter = ui[4] - 0.5 * (sqr(ui[1]) + sqr(ui[2]) + sqr(ui[3])) / ui[0];
    p_inter = p_inter - 0.5 * (sqr(ui[5]) + sqr(ui[6]) + sqr(Bxi));
    const bool fails = (ui[0] <= 0.0) || (p_inter < 0.0);
    bad = bad || ((ev[N + 1] > ev[N]) && fails);
  }
#pragma unroll
  for (int m = 0; m < 7; m++)
    if (roe_nz<N>(m)) f[m] = f[m] - (N == 3 && m == 0 ? coeff[N] : coeff[N] * r[m]);
}

// flux_roe.c:54-259, with the intermediate-state test and HLLE fallback
// (etah = 0, flux_roe.c:40, so the coefficient is |ev|/2 * a).
//
// The fallback (flux_roe.c:233-237) is rare; instead of keeping the four
// input states live through the whole Roe evaluation, it asks `reload` to
// rebuild them (the callers recompute them from memory with the same
// arithmetic, so the values are identical) and evaluates flux_hlle afresh.
// Where the half-sum flux f and the intermediate state's start ui wait while
// the eigen-projection runs: registers by default. A kernel short of registers
// passes LDS slots instead (the values are written once and read back after a
// fence, so they hold no registers in between); the arithmetic is unchanged.
struct RegPark {
  double v[14];
  ATHMI_HDM void put(int q, double x) { v[q] = x; }
  ATHMI_HDM double get(int q) const { return v[q]; }
  ATHMI_HDM void fence() const {}
};
template <class Reload, class Park = RegPark>
ATHMI_HD Cons7 flux_roe_reload(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl,
                               const Prim7& Wr, double Bxi, const EosC& eos, Reload&& reload,
                               Park park = Park()) {
  const RoeAvg r = roe_average(Ul, Ur, Wl, Wr, Bxi);
  RoeSpeeds S;
  roe_speeds(r.droe, r.v1, r.v2, r.v3, r.h, Bxi, r.b2, r.b3, r.x, r.y, eos, true, S);
  const Cons7 Fl = side_flux(Ul, Wl, Bxi, r.pbl);
  const Cons7 Fr = side_flux(Ur, Wr, Bxi, r.pbr);
  if (S.ev0 >= 0.0) return Fl;
  if (S.ev6 <= 0.0) return Fr;
  double f[7];
  {
    const double* fl = &Fl.d;
    const double* fr = &Fr.d;
#pragma unroll
    for (int n = 0; n < 7; n++) f[n] = 0.5 * (fl[n] + fr[n]);
  }
  const double du[7] = {Ur.d - Ul.d, Ur.mx - Ul.mx, Ur.my - Ul.my, Ur.mz - Ul.mz,
                        Ur.e - Ul.e,  Ur.by - Ul.by, Ur.bz - Ul.bz};
  double ui[7] = {Ul.d, Ul.mx, Ul.my, Ul.mz, Ul.e, Ul.by, Ul.bz};
  for (int n = 0; n < 7; n++) {
    park.put(n, f[n]);
    park.put(7 + n, ui[n]);
  }
  RoeVecScalars V;
  roe_vec_scalars(S, r.droe, r.v2, r.v3, eos, V);
  double a[7];
  roe_project(S, V, r.v1, r.v2, r.v3, r.b2, r.b3, r.x, eos, du, a);
  const double ev[7] = {S.ev0, S.ev1, S.ev2, S.ev3, S.ev4, S.ev5, S.ev6};
  double coeff[7];
#pragma unroll
  for (int m = 0; m < 7; m++) coeff[m] = 0.5 * dmax(fabs(ev[m]), 0.0) * a[m];
  park.fence();
  for (int n = 0; n < 7; n++) {
    f[n] = park.get(n);
    ui[n] = park.get(7 + n);
  }
  bool bad = false;
  roe_column_pass<0>(S, V, r.v1, r.v2, r.v3, r.x, eos, a, coeff, ev, Bxi, ui, f, bad);
  roe_column_pass<1>(S, V, r.v1, r.v2, r.v3, r.x, eos, a, coeff, ev, Bxi, ui, f, bad);
  roe_column_pass<2>(S, V, r.v1, r.v2, r.v3, r.x, eos, a, coeff, ev, Bxi, ui, f, bad);
  roe_column_pass<3>(S, V, r.v1, r.v2, r.v3, r.x, eos, a, coeff, ev, Bxi, ui, f, bad);
  roe_column_pass<4>(S, V, r.v1, r.v2, r.v3, r.x, eos, a, coeff, ev, Bxi, ui, f, bad);
  roe_column_pass<5>(S, V, r.v1, r.v2, r.v3, r.x, eos, a, coeff, ev, Bxi, ui, f, bad);
  roe_column_pass<6>(S, V, r.v1, r.v2, r.v3, r.x, eos, a, coeff, ev, Bxi, ui, f, bad);
  if (bad) {
    Cons7 ul, ur;
    Prim7 wl, wr;
    double bx;
    reload(ul, ur, wl, wr, bx);
    return flux_hlle(ul, ur, wl, wr, bx, eos);
  }
  return Cons7{f[0], f[1], f[2], f[3], f[4], f[5], f[6]};
}

enum { kFluxRoe = 0, kFluxHlle = 1 };

// The Riemann solver of a grid, fixed at compile time: a kernel then holds one
// solver only (a runtime choice made both flux paths share one register
// allocation). The Roe solver rebuilds its inputs for the HLLE fallback
// through `reload` and parks f, ui through `park` (see flux_roe_reload).
template <int KIND, class Reload, class Park = RegPark>
ATHMI_HD Cons7 riemann_of(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                          double Bxi, const EosC& eos, Reload&& reload, Park park = Park()) {
  if constexpr (KIND == kFluxHlle) return flux_hlle(Ul, Ur, Wl, Wr, Bxi, eos);
  else return flux_roe_reload(Ul, Ur, Wl, Wr, Bxi, eos, reload, park);
}

}  // namespace athmi
