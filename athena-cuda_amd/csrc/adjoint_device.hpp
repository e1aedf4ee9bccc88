// adjoint_device.hpp -- hand-written reverse (adjoint) of the per-cell and
// per-face routines of the CTU+CT step: conversions, the primitive and Roe
// eigensystems, PLM reconstruction in characteristic variables with CTU
// tracing, the Roe flux (with its HLLE fallback) and the HLLE flux.
//
// Each routine records its forward pass in the reference's evaluation order
// (the oracle's order, so that under -ffp-contract=off the recomputation is
// bit-identical to the forward kernels and every branch -- limiter, clamps,
// upwinding, HLLE fallback, alpha and beta degeneracies -- is the one the
// forward pass took), then propagates adjoints statement by statement in
// reverse. Derivatives are those of the branch taken; sqrt at 0 and the max/
// min ties contribute zero (the convention of the tangent-linear model).
//
// The same source is compiled for the device (adjoint kernels) and for the
// host (athmi_vjp_* entry points used to check it).
#pragma once
#include "mhd_device.hpp"

namespace athmi {
namespace adj {


ATHMI_HD double sgnabs(double x) { return (x < 0.0) ? -1.0 : 1.0; }  // d|x|/dx

// a / b in derivative arithmetic. The primal values a reverse recomputes
// (they decide its branches) keep the correctly rounded division; derivative
// values need accuracy only. On the device: the hardware reciprocal estimate
// (relative error below 2^-24), one Newton-Raphson step r1 = r0 + r0 (1 - b r0)
// and one residual correction of the quotient q0 + (a - b q0) r1, both with
// fused multiply-adds: six instructions against the ten of the correctly
// rounded sequence, and within an ulp of it (equal to it on 2^20 random
// quotients measured; -0 / b gives +0). On the host: a / b.
ATHMI_HD double adiv(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r0 = __builtin_amdgcn_rcp(b);
  const double r1 = fma(r0, fma(-b, r0, 1.0), r0);
  const double q0 = a * r1;
  return fma(fma(-b, q0, a), r1, q0);
#else
  return a / b;
#endif
}

// adjoint of y = sqrt(x): xb += yb / (2 y), zero at y == 0
ATHMI_HD double sqrt_b(double y, double yb) { return (y > 0.0) ? adiv(0.5 * yb, y) : 0.0; }

ATHMI_HD double* arr(Cons7& c) { return &c.d; }
ATHMI_HD const double* arr(const Cons7& c) { return &c.d; }
ATHMI_HD double* arr(Prim7& c) { return &c.d; }
ATHMI_HD const double* arr(const Prim7& c) { return &c.d; }
ATHMI_HD Cons7 zero_cons() { return Cons7{0, 0, 0, 0, 0, 0, 0}; }
ATHMI_HD Prim7 zero_prim() { return Prim7{0, 0, 0, 0, 0, 0, 0}; }

// ---------------------------------------------------------------------------
// convert_var.c:37-55 reversed: wb -> ub (+=), bxb (+=)
ATHMI_HD void cons_to_prim_b(const Cons7& u, double bx, const EosC& eos, const Prim7& wb,
                             Cons7& ub, double& bxb) {
  const double di = 1.0 / u.d;
  const double ke = 0.5 * (sqr(u.mx) + sqr(u.my) + sqr(u.mz));
  double p = u.e - ke * di;
  p = p - 0.5 * (sqr(bx) + sqr(u.by) + sqr(u.bz));
  p = p * eos.g1;
  const double p2b = (p > kTiny) ? wb.p : 0.0;  // MAX(p, TINY_NUMBER)
  const double p1b = p2b * eos.g1;
  bxb -= p1b * bx;
  ub.by += wb.by - p1b * u.by;
  ub.bz += wb.bz - p1b * u.bz;
  ub.e += p1b;
  const double keb = -p1b * di;
  double dib = -p1b * ke;
  ub.mx += keb * u.mx + wb.vx * di;
  ub.my += keb * u.my + wb.vy * di;
  ub.mz += keb * u.mz + wb.vz * di;
  dib += wb.vx * u.mx + wb.vy * u.my + wb.vz * u.mz;
  ub.d += wb.d - dib * di * di;
}

// convert_var.c:64-92 reversed
ATHMI_HD void prim_to_cons_b(const Prim7& w, double bx, const EosC& eos, const Cons7& ub,
                             Prim7& wb, double& bxb) {
  const double v2 = sqr(w.vx) + sqr(w.vy) + sqr(w.vz);
  wb.d += ub.d + ub.mx * w.vx + ub.my * w.vy + ub.mz * w.vz + ub.e * 0.5 * v2;
  wb.vx += ub.mx * w.d + ub.e * w.d * w.vx;
  wb.vy += ub.my * w.d + ub.e * w.d * w.vy;
  wb.vz += ub.mz * w.d + ub.e * w.d * w.vz;
  wb.p += adiv(ub.e, eos.g1);
  bxb += ub.e * bx;
  wb.by += ub.by + ub.e * w.by;
  wb.bz += ub.bz + ub.e * w.bz;
}

// ---------------------------------------------------------------------------
// esystem_prim.c:36-235: forward record. R(i,j) / L(i,j) give the entries
// (structural zeros as 0) in the reference's expressions.
struct EsysPrimF {
  double d, v1, p, b1, b2, b3, gamma;
  double di, btsq, vaxsq, asq, ct2, tsum, tdif, cf2_cs2, cfsq, cf, cssq, cs, bt, bet2, bet3;
  double alpha_f, alpha_s, sqrtd, s, a, qf, qs, af, as, vax;
  double na, qfn, qsn, afp, asp;
  int abr;  // alpha branch: 0 degenerate constants, 1 general
  int btz;  // bt == 0
  double ev[7];

  ATHMI_HDM void run(double d_, double v1_, double p_, double b1_, double b2_, double b3_,
                    double gamma_) {
    d = d_; v1 = v1_; p = p_; b1 = b1_; b2 = b2_; b3 = b3_; gamma = gamma_;
    di = 1.0 / d;
    btsq = b2 * b2 + b3 * b3;
    vaxsq = b1 * b1 * di;
    asq = gamma * p * di;
    ct2 = btsq * di;
    tsum = vaxsq + ct2 + asq;
    tdif = vaxsq + ct2 - asq;
    cf2_cs2 = sqrt(tdif * tdif + 4.0 * asq * ct2);
    cfsq = 0.5 * (tsum + cf2_cs2);
    cf = sqrt(cfsq);
    cssq = asq * vaxsq / cfsq;
    cs = sqrt(cssq);
    bt = sqrt(btsq);
    btz = (bt == 0.0);
    if (btz) {
      bet2 = 1.0;
      bet3 = 0.0;
    } else {
      bet2 = b2 / bt;
      bet3 = b3 / bt;
    }
    abr = 0;
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
      abr = 1;
      alpha_f = sqrt((asq - cssq) / cf2_cs2);
      alpha_s = sqrt((cfsq - asq) / cf2_cs2);
    }
    sqrtd = sqrt(d);
    s = dsign(b1);
    a = sqrt(asq);
    qf = cf * alpha_f * s;
    qs = cs * alpha_s * s;
    af = a * alpha_f * sqrtd;
    as = a * alpha_s * sqrtd;
    vax = sqrt(vaxsq);
    ev[0] = v1 - cf;
    ev[1] = v1 - vax;
    ev[2] = v1 - cs;
    ev[3] = v1;
    ev[4] = v1 + cs;
    ev[5] = v1 + vax;
    ev[6] = v1 + cf;
    na = 0.5 / asq;
    qfn = na * qf;
    qsn = na * qs;
    afp = na * af * di;
    asp = na * as * di;
  }

  ATHMI_HDM double R(int i, int j) const {
    switch (i * 7 + j) {
      case 0: case 6: return d * alpha_f;
      case 2: case 4: return d * alpha_s;
      case 3: return 1.0;
      case 7: return -cf * alpha_f;
      case 13: return cf * alpha_f;
      case 9: return -cs * alpha_s;
      case 11: return cs * alpha_s;
      case 14: return qs * bet2;
      case 20: return -(qs * bet2);
      case 15: return -bet3;
      case 19: return bet3;
      case 16: return -qf * bet2;
      case 18: return qf * bet2;
      case 21: return qs * bet3;
      case 27: return -(qs * bet3);
      case 22: return bet2;
      case 26: return -bet2;
      case 23: return -qf * bet3;
      case 25: return qf * bet3;
      case 28: case 34: return d * asq * alpha_f;
      case 30: case 32: return d * asq * alpha_s;
      case 35: case 41: return as * bet2;
      case 36: case 40: return -bet3 * s * sqrtd;
      case 37: case 39: return -af * bet2;
      case 42: case 48: return as * bet3;
      case 43: case 47: return bet2 * s * sqrtd;
      case 44: case 46: return -af * bet3;
      default: return 0.0;
    }
  }
  ATHMI_HDM double L(int i, int j) const {
    switch (i * 7 + j) {
      case 1: return -na * cf * alpha_f;
      case 43: return -(-na * cf * alpha_f);
      case 2: return qsn * bet2;
      case 44: return -(qsn * bet2);
      case 3: return qsn * bet3;
      case 45: return -(qsn * bet3);
      case 4: case 46: return na * alpha_f * di;
      case 5: case 47: return asp * bet2;
      case 6: case 48: return asp * bet3;
      case 9: return -0.5 * bet3;
      case 37: return -(-0.5 * bet3);
      case 10: return 0.5 * bet2;
      case 38: return -(0.5 * bet2);
      case 12: case 40: return -0.5 * bet3 * s / sqrtd;
      case 13: case 41: return 0.5 * bet2 * s / sqrtd;
      case 15: return -na * cs * alpha_s;
      case 29: return -(-na * cs * alpha_s);
      case 16: return -qfn * bet2;
      case 30: return -(-qfn * bet2);
      case 17: return -qfn * bet3;
      case 31: return -(-qfn * bet3);
      case 18: case 32: return na * alpha_s * di;
      case 19: case 33: return -afp * bet2;
      case 20: case 34: return -afp * bet3;
      case 21: return 1.0;
      case 25: return -1.0 / asq;
      default: return 0.0;
    }
  }

  // Adjoints of the intermediates the entries of R and L are built from;
  // each entry's adjoint is folded in as it arrives (addR, addL), so no
  // 7x7 adjoint matrix is ever held.
  struct Acc {
    double db = 0, alpha_fb = 0, alpha_sb = 0, cfb = 0, csb = 0, qsb = 0, qfb = 0, bet2b = 0;
    double bet3b = 0, tb = 0, asqb = 0, asb = 0, afb = 0, sqrtdb = 0, nab = 0, qsnb = 0;
    double qfnb = 0, aspb = 0, afpb = 0, dib = 0;
  };
  // A += X * dR(i,j)/d(intermediates); R entries as in R(i, j)
  ATHMI_HDM void addR(int i, int j, double X, Acc& A) const {
    switch (i * 7 + j) {
      case 0: case 6: A.db += X * alpha_f; A.alpha_fb += X * d; break;
      case 2: case 4: A.db += X * alpha_s; A.alpha_sb += X * d; break;
      case 7: A.cfb -= X * alpha_f; A.alpha_fb -= X * cf; break;
      case 13: A.cfb += X * alpha_f; A.alpha_fb += X * cf; break;
      case 9: A.csb -= X * alpha_s; A.alpha_sb -= X * cs; break;
      case 11: A.csb += X * alpha_s; A.alpha_sb += X * cs; break;
      case 14: A.qsb += X * bet2; A.bet2b += X * qs; break;
      case 20: A.qsb -= X * bet2; A.bet2b -= X * qs; break;
      case 15: A.bet3b -= X; break;
      case 19: A.bet3b += X; break;
      case 16: A.qfb -= X * bet2; A.bet2b -= X * qf; break;
      case 18: A.qfb += X * bet2; A.bet2b += X * qf; break;
      case 21: A.qsb += X * bet3; A.bet3b += X * qs; break;
      case 27: A.qsb -= X * bet3; A.bet3b -= X * qs; break;
      case 22: A.bet2b += X; break;
      case 26: A.bet2b -= X; break;
      case 23: A.qfb -= X * bet3; A.bet3b -= X * qf; break;
      case 25: A.qfb += X * bet3; A.bet3b += X * qf; break;
      case 28: case 34: A.alpha_fb += X * d * asq; A.tb += X * alpha_f; break;
      case 30: case 32: A.alpha_sb += X * d * asq; A.tb += X * alpha_s; break;
      case 35: case 41: A.asb += X * bet2; A.bet2b += X * as; break;
      case 36: case 40: A.bet3b -= X * s * sqrtd; A.sqrtdb -= X * bet3 * s; break;
      case 37: case 39: A.afb -= X * bet2; A.bet2b -= X * af; break;
      case 42: case 48: A.asb += X * bet3; A.bet3b += X * as; break;
      case 43: case 47: A.bet2b += X * s * sqrtd; A.sqrtdb += X * bet2 * s; break;
      case 44: case 46: A.afb -= X * bet3; A.bet3b -= X * af; break;
      default: break;
    }
  }
  // A += X * dL(i,j)/d(intermediates); L entries as in L(i, j)
  ATHMI_HDM void addL(int i, int j, double X, Acc& A) const {
    switch (i * 7 + j) {
      case 1: case 43: {
        const double Y = (i == 0) ? X : -X;  // L61 = -L01 = na cf alpha_f
        A.nab -= Y * cf * alpha_f; A.cfb -= Y * na * alpha_f; A.alpha_fb -= Y * na * cf;
        break;
      }
      case 2: A.qsnb += X * bet2; A.bet2b += X * qsn; break;
      case 44: A.qsnb -= X * bet2; A.bet2b -= X * qsn; break;
      case 3: A.qsnb += X * bet3; A.bet3b += X * qsn; break;
      case 45: A.qsnb -= X * bet3; A.bet3b -= X * qsn; break;
      case 4: case 46: A.nab += X * alpha_f * di; A.alpha_fb += X * na * di; A.dib += X * na * alpha_f; break;
      case 5: case 47: A.aspb += X * bet2; A.bet2b += X * asp; break;
      case 6: case 48: A.aspb += X * bet3; A.bet3b += X * asp; break;
      case 9: A.bet3b -= 0.5 * X; break;
      case 37: A.bet3b += 0.5 * X; break;
      case 10: A.bet2b += 0.5 * X; break;
      case 38: A.bet2b -= 0.5 * X; break;
      case 12: case 40:
        A.bet3b -= adiv(0.5 * s * X, sqrtd); A.sqrtdb += adiv(X * 0.5 * bet3 * s, sqrtd * sqrtd); break;
      case 13: case 41:
        A.bet2b += adiv(0.5 * s * X, sqrtd); A.sqrtdb -= adiv(X * 0.5 * bet2 * s, sqrtd * sqrtd); break;
      case 15: case 29: {
        const double Y = (i == 2) ? X : -X;  // L41 = -L21
        A.nab -= Y * cs * alpha_s; A.csb -= Y * na * alpha_s; A.alpha_sb -= Y * na * cs;
        break;
      }
      case 16: A.qfnb -= X * bet2; A.bet2b -= X * qfn; break;
      case 30: A.qfnb += X * bet2; A.bet2b += X * qfn; break;
      case 17: A.qfnb -= X * bet3; A.bet3b -= X * qfn; break;
      case 31: A.qfnb += X * bet3; A.bet3b += X * qfn; break;
      case 18: case 32: A.nab += X * alpha_s * di; A.alpha_sb += X * na * di; A.dib += X * na * alpha_s; break;
      case 19: case 33: A.afpb -= X * bet2; A.bet2b -= X * afp; break;
      case 20: case 34: A.afpb -= X * bet3; A.bet3b -= X * afp; break;
      case 25: A.asqb += adiv(X, asq * asq); break;
      default: break;
    }
  }

  // Adjoint from the folded entry adjoints A and evb[7]: accumulates the
  // adjoints of (d, v1, p, b1, b2, b3) into in_b[6].
  ATHMI_HDM void back_acc(const Acc& A0, const double* evb, double* in_b) const {
    Acc A = A0;
    double v1b = 0, pb = 0, b1b = 0, b2b = 0, b3b = 0, vaxb = 0;
    A.db += A.tb * asq;
    A.asqb += A.tb * d;
    // L prefixes: afp = na*af*di, asp = na*as*di, qfn = na*qf, qsn = na*qs, na = 0.5/asq
    A.nab += A.afpb * af * di + A.aspb * as * di + A.qfnb * qf + A.qsnb * qs;
    A.afb += A.afpb * na * di;
    A.asb += A.aspb * na * di;
    A.dib += A.afpb * na * af + A.aspb * na * as;
    A.qfb += A.qfnb * na;
    A.qsb += A.qsnb * na;
    A.asqb -= adiv(A.nab * na, asq);
    // eigenvalues
    for (int k = 0; k < 7; k++) v1b += evb[k];
    A.cfb += evb[6] - evb[0];
    vaxb += evb[5] - evb[1];
    A.csb += evb[4] - evb[2];
    double vaxsqb = sqrt_b(vax, vaxb);
    // af = a*alpha_f*sqrtd, as = a*alpha_s*sqrtd
    const double ab = A.afb * alpha_f * sqrtd + A.asb * alpha_s * sqrtd;
    A.alpha_fb += A.afb * a * sqrtd;
    A.alpha_sb += A.asb * a * sqrtd;
    A.sqrtdb += A.afb * a * alpha_f + A.asb * a * alpha_s;
    // qf = cf*alpha_f*s, qs = cs*alpha_s*s
    A.cfb += A.qfb * alpha_f * s;
    A.alpha_fb += A.qfb * cf * s;
    A.csb += A.qsb * alpha_s * s;
    A.alpha_sb += A.qsb * cs * s;
    A.asqb += sqrt_b(a, ab);
    A.db += sqrt_b(sqrtd, A.sqrtdb);
    double cssqb = 0, cfsqb = 0, cf2_cs2b = 0;
    if (abr) {
      const double Qf = (asq - cssq) / cf2_cs2, Qs = (cfsq - asq) / cf2_cs2;
      const double Qfb = sqrt_b(alpha_f, A.alpha_fb), Qsb = sqrt_b(alpha_s, A.alpha_sb);
      A.asqb += adiv(Qfb - Qsb, cf2_cs2);
      cssqb -= adiv(Qfb, cf2_cs2);
      cfsqb += adiv(Qsb, cf2_cs2);
      cf2_cs2b -= adiv(Qfb * Qf + Qsb * Qs, cf2_cs2);
    }
    double btsqb = 0;
    if (!btz) {
      b2b += adiv(A.bet2b, bt);
      b3b += adiv(A.bet3b, bt);
      const double btb = -adiv(A.bet2b * bet2 + A.bet3b * bet3, bt);
      btsqb += sqrt_b(bt, btb);
    }
    cssqb += sqrt_b(cs, A.csb);
    A.asqb += adiv(cssqb * vaxsq, cfsq);
    vaxsqb += adiv(cssqb * asq, cfsq);
    cfsqb -= adiv(cssqb * cssq, cfsq);
    cfsqb += sqrt_b(cf, A.cfb);
    const double tsumb = 0.5 * cfsqb;
    cf2_cs2b += 0.5 * cfsqb;
    const double Yb = sqrt_b(cf2_cs2, cf2_cs2b);
    const double tdifb = 2.0 * tdif * Yb;
    A.asqb += 4.0 * ct2 * Yb;
    double ct2b = 4.0 * asq * Yb;
    vaxsqb += tsumb + tdifb;
    ct2b += tsumb + tdifb;
    A.asqb += tsumb - tdifb;
    btsqb += ct2b * di;
    A.dib += ct2b * btsq;
    pb += A.asqb * gamma * di;
    A.dib += A.asqb * gamma * p;
    b1b += vaxsqb * 2.0 * b1 * di;
    A.dib += vaxsqb * b1 * b1;
    b2b += 2.0 * b2 * btsqb;
    b3b += 2.0 * b3 * btsqb;
    A.db -= A.dib * di * di;
    in_b[0] += A.db;
    in_b[1] += v1b;
    in_b[2] += pb;
    in_b[3] += b1b;
    in_b[4] += b2b;
    in_b[5] += b3b;
  }

  // Adjoint from matrices of entry adjoints Rb(i,j), Lb(i,j) and evb[7]
  template <class RB, class LB>
  ATHMI_HDM void back(const RB& Rb, const LB& Lb, const double* evb, double* in_b) const {
    Acc A;
#pragma unroll
    for (int i = 0; i < 7; i++)
#pragma unroll
      for (int j = 0; j < 7; j++) {
        addR(i, j, Rb(i, j), A);
        addL(i, j, Lb(i, j), A);
      }
    back_acc(A, evb, in_b);
  }
};

// never-written entries of the primitive eigenmatrices (esystem_prim.c:116-234);
// skipping them in a sum changes no finite value
ATHMI_HD constexpr bool prim_Lnz(int n, int m) {
  return (n == 0 || n == 2 || n == 4 || n == 6) ? (m != 0)
         : (n == 1 || n == 5) ? (m == 2 || m == 3 || m == 5 || m == 6)
                              : (m == 0 || m == 4);
}
ATHMI_HD constexpr bool prim_Rnz(int n, int m) {
  return (n == 0) ? (m != 1 && m != 5)
         : (n == 1 || n == 4) ? (m == 0 || m == 2 || m == 4 || m == 6)
                              : (m != 3);
}

// After this point values held in memory are read again: a structure kept
// in LDS is then not cached in registers from one reverse stage to the next.
ATHMI_HD void reload_fence() { asm volatile("" ::: "memory"); }

// ---------------------------------------------------------------------------
// lr_states_plm.c:64-248 for one cell i (W[i-1], W[i], W[i+1], Bxc[i]):
// out_l -> Wl[i+1], out_r -> Wr[i]. Only the stencil and the eigensystem are
// kept; the forward chain (Steps 2-7) is recomputed, in the reference order,
// by each reverse stage that needs a part of it, which keeps few values live.
// PlmCore holds what the reverse reads (it may live in LDS, one per thread);
// PlmF adds the forward outputs.
struct PlmCore {
  double wm[7], w0[7], wp[7], bxc, dtodx;
  EsysPrimF E;
  // quotients every stage that re-runs a part of the forward chain needs,
  // computed once (the same expressions, so the same values): the limiter's
  // harmonic means 2 l r / (l + r) per component, and the entries of L that
  // divide, L[1][5] = L[5][5], L[1][6] = L[5][6] and L[3][4]
  double gm[7], lq[3];

  // the stencil and the eigensystem of W[i] (Step 1)
  ATHMI_HDM void prepare(const Prim7& Wm, const Prim7& W0, const Prim7& Wp, double bxc_,
                         double dtodx_, const EosC& eos) {
    for (int n = 0; n < 7; n++) {
      wm[n] = arr(Wm)[n];
      w0[n] = arr(W0)[n];
      wp[n] = arr(Wp)[n];
    }
    bxc = bxc_;
    dtodx = dtodx_;
    E.run(W0.d, W0.vx, W0.p, bxc, W0.by, W0.bz, eos.gamma);
    for (int m = 0; m < 7; m++) {
      const double l = w0[m] - wm[m], r = wp[m] - w0[m];
      gm[m] = (l * r > 0.0) ? 2.0 * l * r / (l + r) : 0.0;
    }
    lq[0] = E.L(1, 5);
    lq[1] = E.L(1, 6);
    lq[2] = E.L(3, 4);
  }
  // entry (n, m) of L
  ATHMI_HDM double Lq(int n, int m) const {
    if ((n == 1 || n == 5) && m == 5) return lq[0];
    if ((n == 1 || n == 5) && m == 6) return lq[1];
    if (n == 3 && m == 4) return lq[2];
    return E.L(n, m);
  }

  // Steps 2-4 for component n: the four characteristic projections and the
  // limited slope
  ATHMI_HDM void diffs(int m, double& c, double& l, double& r, double& g) const {
    c = wp[m] - wm[m];
    l = w0[m] - wm[m];
    r = wp[m] - w0[m];
    g = gm[m];
  }
  ATHMI_HDM void project(int n, double& ac, double& al, double& ar, double& ag) const {
    double c = 0.0, l = 0.0, r = 0.0, g = 0.0;
#pragma unroll
    for (int m = 0; m < 7; m++) {
      if (!prim_Lnz(n, m)) continue;
      double dc, dl, dr, dg;
      diffs(m, dc, dl, dr, dg);
      const double L = Lq(n, m);
      c = c + L * dc;
      l = l + L * dl;
      r = r + L * dr;
      g = g + L * dg;
    }
    ac = c;
    al = l;
    ar = r;
    ag = g;
  }
  ATHMI_HDM double limited(double ac, double al, double ar, double ag) const {
    if (!(al * ar > 0.0)) return 0.0;
    const double ls1 = dmin(fabs(al), fabs(ar));
    const double ls2 = dmin(0.5 * fabs(ac), fabs(ag));
    return dsign(ac) * dmin(2.0 * ls1, ls2);
  }
  // Steps 2-7: slopes da, the unclamped Wlv0/Wrv0, the clamped Wlv/Wrv and
  // dW = Wrv - Wlv (any output may be null)
  ATHMI_HDM void chain(double* da_out, double* Wlv0, double* Wrv0, double* Wlv, double* Wrv,
                       double* dW) const {
    double da[7];
#pragma unroll
    for (int n = 0; n < 7; n++) {
      double ac, al, ar, ag;
      project(n, ac, al, ar, ag);
      da[n] = limited(ac, al, ar, ag);
      if (da_out) da_out[n] = da[n];
    }
#pragma unroll
    for (int n = 0; n < 7; n++) {
      double acc = 0.0;
#pragma unroll
      for (int m = 0; m < 7; m++)
        if (prim_Rnz(n, m)) acc = acc + da[m] * E.R(n, m);
      const double l0 = w0[n] - 0.5 * acc, r0 = w0[n] + 0.5 * acc;
      const double l = dmin(dmax(w0[n], wm[n]), dmax(dmin(w0[n], wm[n]), l0));
      const double r = dmin(dmax(w0[n], wp[n]), dmax(dmin(w0[n], wp[n]), r0));
      if (Wlv0) Wlv0[n] = l0;
      if (Wrv0) Wrv0[n] = r0;
      if (Wlv) Wlv[n] = l;
      if (Wrv) Wrv[n] = r;
      if (dW) dW[n] = r - l;
    }
  }

  // Adjoint: out_lb, out_rb -> (+=) wmb, w0b, wpb, bxcb. The adjoints of the
  // eigenmatrix entries are folded into the eigensystem's intermediates as
  // they arise (EsysPrimF::addR / addL).
  ATHMI_HDM void back(const double* out_lb, const double* out_rb, double* wmb, double* w0b,
                      double* wpb, double& bxcb) const {
    EsysPrimF::Acc A;
    double evb[7], dWb[7];
    for (int n = 0; n < 7; n++) {
      evb[n] = 0.0;
      dWb[n] = 0.0;
    }
    const double c = 0.5 * dtodx;
    {
      double dW[7];
      chain(nullptr, nullptr, nullptr, nullptr, nullptr, dW);
      // Step 9
#pragma unroll
      for (int n = 0; n < 7; n++) {
        reload_fence();
        if (!(E.ev[n] > 0.) && !(E.ev[n] < 0.)) continue;
        double LdW = 0.0;
#pragma unroll
        for (int m = 0; m < 7; m++)
          if (prim_Lnz(n, m)) LdW += Lq(n, m) * dW[m];
        if (E.ev[n] > 0.) {  // out_l += qa R[:,n], qa = c (ev6 - evn) LdW
          double qab = 0.0;
#pragma unroll
          for (int m = 0; m < 7; m++)
            if (prim_Rnz(m, n)) qab += out_lb[m] * E.R(m, n);
          const double f = c * (E.ev[6] - E.ev[n]);
          const double qa = f * LdW;
#pragma unroll
          for (int m = 0; m < 7; m++) {
            if (prim_Rnz(m, n)) E.addR(m, n, out_lb[m] * qa, A);
            if (prim_Lnz(n, m)) {
              E.addL(n, m, qab * f * dW[m], A);
              dWb[m] += qab * f * Lq(n, m);
            }
          }
          evb[6] += qab * c * LdW;
          evb[n] -= qab * c * LdW;
        } else {
          // out_r += qa R[:,n], qa = c (ev0 - evn) LdW; out_l -= qa2 R[:,n]
          double qab = 0.0, qa2b = 0.0;
#pragma unroll
          for (int m = 0; m < 7; m++)
            if (prim_Rnz(m, n)) {
              qab += out_rb[m] * E.R(m, n);
              qa2b -= out_lb[m] * E.R(m, n);
            }
          const double f = c * (E.ev[0] - E.ev[n]);
          const double f2 = c * (E.ev[n] - E.ev[6]);
          const double k = qab * f + qa2b * f2;
#pragma unroll
          for (int m = 0; m < 7; m++) {
            if (prim_Rnz(m, n)) E.addR(m, n, out_rb[m] * (f * LdW) - out_lb[m] * (f2 * LdW), A);
            if (prim_Lnz(n, m)) {
              E.addL(n, m, k * dW[m], A);
              dWb[m] += k * Lq(n, m);
            }
          }
          evb[0] += qab * c * LdW;
          evb[n] += -qab * c * LdW + qa2b * c * LdW;
          evb[6] -= qa2b * c * LdW;
        }
      }
      // Step 8
      reload_fence();
      const double qx1 = 0.5 * dmax(E.ev[6], 0.0) * dtodx;
      const double qx2 = -0.5 * dmin(E.ev[0], 0.0) * dtodx;
      double qx1b = 0.0, qx2b = 0.0;
#pragma unroll
      for (int m = 0; m < 7; m++) {
        qx1b -= out_lb[m] * dW[m];
        qx2b += out_rb[m] * dW[m];
        dWb[m] += qx2 * out_rb[m] - qx1 * out_lb[m];
      }
      if (E.ev[6] > 0.0) evb[6] += qx1b * 0.5 * dtodx;
      if (E.ev[0] < 0.0) evb[0] -= qx2b * 0.5 * dtodx;
    }
    // Step 7: clamps, then Step 5 (dWm = R da)
    {
      double da[7], Wlv0[7], Wrv0[7];
      reload_fence();
      chain(da, Wlv0, Wrv0, nullptr, nullptr, nullptr);
      double dWmb[7];
#pragma unroll
      for (int n = 0; n < 7; n++) {
        const double Wrvb = out_lb[n] + dWb[n];
        const double Wlvb = out_rb[n] - dWb[n];
        // Wlv = MIN(MAX(w0,wm), MAX(MIN(w0,wm), Wlv0))
        double lob = 0, hib = 0, s1b = 0, Wlv0b = 0, Wrv0b = 0;
        {
          const double lo = dmin(w0[n], wm[n]), hi = dmax(w0[n], wm[n]);
          const double s1 = dmax(lo, Wlv0[n]);
          if (hi < s1) hib = Wlvb; else s1b = Wlvb;
          if (lo > Wlv0[n]) lob = s1b; else Wlv0b = s1b;
          if (w0[n] < wm[n]) w0b[n] += lob; else wmb[n] += lob;
          if (w0[n] > wm[n]) w0b[n] += hib; else wmb[n] += hib;
        }
        {
          lob = hib = s1b = 0;
          const double lo = dmin(w0[n], wp[n]), hi = dmax(w0[n], wp[n]);
          const double s1 = dmax(lo, Wrv0[n]);
          if (hi < s1) hib = Wrvb; else s1b = Wrvb;
          if (lo > Wrv0[n]) lob = s1b; else Wrv0b = s1b;
          if (w0[n] < wp[n]) w0b[n] += lob; else wpb[n] += lob;
          if (w0[n] > wp[n]) w0b[n] += hib; else wpb[n] += hib;
        }
        w0b[n] += Wlv0b + Wrv0b;
        dWmb[n] = 0.5 * (Wrv0b - Wlv0b);
      }
#pragma unroll
      for (int m = 0; m < 7; m++) {
        double acc = 0.0;
#pragma unroll
        for (int n = 0; n < 7; n++)
          if (prim_Rnz(n, m)) {
            acc += dWmb[n] * E.R(n, m);
            E.addR(n, m, dWmb[n] * da[m], A);
          }
        dWb[m] = acc;  // reused: the adjoint of the slope da[m]
      }
    }
    // Step 4 (limiter) and Step 3 (projections) per characteristic field n
    double dWcb[7], dWlb[7], dWrb[7], dWgb[7];
    for (int m = 0; m < 7; m++) dWcb[m] = dWlb[m] = dWrb[m] = dWgb[m] = 0.0;
#pragma unroll
    for (int n = 0; n < 7; n++) {
      double ac, al, ar, ag;
      project(n, ac, al, ar, ag);
      double acb = 0.0, alb = 0.0, arb = 0.0, agb = 0.0;
      if (al * ar > 0.0) {
        const double ls1 = dmin(fabs(al), fabs(ar));
        const double ls2 = dmin(0.5 * fabs(ac), fabs(ag));
        const double sg = dsign(ac);
        double ls1b = 0.0, ls2b = 0.0;
        if (2.0 * ls1 < ls2) ls1b = 2.0 * sg * dWb[n]; else ls2b = sg * dWb[n];
        if (fabs(al) < fabs(ar)) alb += ls1b * sgnabs(al);
        else arb += ls1b * sgnabs(ar);
        if (0.5 * fabs(ac) < fabs(ag)) acb += ls2b * 0.5 * sgnabs(ac);
        else agb += ls2b * sgnabs(ag);
      }
#pragma unroll
      for (int m = 0; m < 7; m++) {
        if (!prim_Lnz(n, m)) continue;
        double dc, dl, dr, dg;
        diffs(m, dc, dl, dr, dg);
        E.addL(n, m, acb * dc + alb * dl + arb * dr + agb * dg, A);
        const double L = Lq(n, m);
        dWcb[m] += acb * L;
        dWlb[m] += alb * L;
        dWrb[m] += arb * L;
        dWgb[m] += agb * L;
      }
    }
    // Step 2
#pragma unroll
    for (int m = 0; m < 7; m++) {
      const double dl = w0[m] - wm[m], dr = wp[m] - w0[m];
      double lb = dWlb[m], rb = dWrb[m];
      if (dl * dr > 0.0) {
        const double S = dl + dr;
        lb += adiv(dWgb[m] * 2.0 * dr * dr, S * S);
        rb += adiv(dWgb[m] * 2.0 * dl * dl, S * S);
      }
      wpb[m] += dWcb[m] + rb;
      wmb[m] -= dWcb[m] + lb;
      w0b[m] += lb - rb;
    }
    // eigensystem of W[i] (d, vx, p, Bxc, by, bz)
    double inb[6] = {0, 0, 0, 0, 0, 0};
    reload_fence();
    E.back_acc(A, evb, inb);
    w0b[0] += inb[0];
    w0b[1] += inb[1];
    w0b[4] += inb[2];
    bxcb += inb[3];
    w0b[5] += inb[4];
    w0b[6] += inb[5];
  }
};

struct PlmF : PlmCore {
  double out_l[7], out_r[7];

  ATHMI_HDM void run(const Prim7& Wm, const Prim7& W0, const Prim7& Wp, double bxc_,
                     double dtodx_, const EosC& eos) {
    prepare(Wm, W0, Wp, bxc_, dtodx_, eos);
    double Wlv[7], Wrv[7], dW[7];
    chain(nullptr, nullptr, nullptr, Wlv, Wrv, dW);
    const double qx1 = 0.5 * dmax(E.ev[6], 0.0) * dtodx;  // Step 8 (:181-192)
    for (int n = 0; n < 7; n++) out_l[n] = Wrv[n] - qx1 * dW[n];
    const double qx2 = -0.5 * dmin(E.ev[0], 0.0) * dtodx;
    for (int n = 0; n < 7; n++) out_r[n] = Wlv[n] + qx2 * dW[n];
    for (int n = 0; n < 7; n++) {  // Step 9 (:201-209)
      if (E.ev[n] > 0.) {
        double qa = 0.0;
        for (int m = 0; m < 7; m++)
          if (prim_Lnz(n, m)) qa = qa + Lq(n, m) * 0.5 * dtodx * (E.ev[6] - E.ev[n]) * dW[m];
        for (int m = 0; m < 7; m++)
          if (prim_Rnz(m, n)) out_l[m] = out_l[m] + qa * E.R(m, n);
      }
    }
    for (int n = 0; n < 7; n++) {  // (:211-224)
      if (E.ev[n] < 0.) {
        double qa = 0.0;
        for (int m = 0; m < 7; m++)
          if (prim_Lnz(n, m)) qa = qa + Lq(n, m) * 0.5 * dtodx * (E.ev[0] - E.ev[n]) * dW[m];
        for (int m = 0; m < 7; m++)
          if (prim_Rnz(m, n)) out_r[m] = out_r[m] + qa * E.R(m, n);
        qa = 0.0;
        for (int m = 0; m < 7; m++)
          if (prim_Lnz(n, m)) qa = qa + Lq(n, m) * 0.5 * dtodx * (E.ev[n] - E.ev[6]) * dW[m];
        for (int m = 0; m < 7; m++)
          if (prim_Rnz(m, n)) out_l[m] = out_l[m] - qa * E.R(m, n);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Roe averages (flux_roe.c:91-114, flux_hlle.c:87-110)
struct RoeAvgF {
  double sdl, sdr, isd, droe, v1, v2, v3, b2, b3, x, y, pbl, pbr, h;
  ATHMI_HDM void run(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                    double Bxi) {
    sdl = sqrt(Wl.d);
    sdr = sqrt(Wr.d);
    isd = 1.0 / (sdl + sdr);
    droe = sdl * sdr;
    v1 = (sdl * Wl.vx + sdr * Wr.vx) * isd;
    v2 = (sdl * Wl.vy + sdr * Wr.vy) * isd;
    v3 = (sdl * Wl.vz + sdr * Wr.vz) * isd;
    b2 = (sdr * Wl.by + sdl * Wr.by) * isd;
    b3 = (sdr * Wl.bz + sdl * Wr.bz) * isd;
    x = 0.5 * (sqr(Wl.by - Wr.by) + sqr(Wl.bz - Wr.bz)) / (sqr(sdl + sdr));
    y = 0.5 * (Wl.d + Wr.d) / droe;
    pbl = 0.5 * (sqr(Bxi) + sqr(Wl.by) + sqr(Wl.bz));
    pbr = 0.5 * (sqr(Bxi) + sqr(Wr.by) + sqr(Wr.bz));
    h = ((Ul.e + Wl.p + pbl) / sdl + (Ur.e + Wr.p + pbr) / sdr) * isd;
  }
  // b: adjoints of (droe, v1, v2, v3, h, b2, b3, x, y) in the esys_roe input
  // order (d, v1, v2, v3, h, [b1 = Bxi], b2, b3, x, y) -> b[0..9] with b[5] the
  // Bxi adjoint; pblb/pbrb: adjoints of pbl, pbr from the side fluxes.
  ATHMI_HDM void back(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                     double Bxi, const double* b, double pblb, double pbrb, Cons7& Ulb,
                     Cons7& Urb, Prim7& Wlb, Prim7& Wrb, double& Bxib) const {
    double sdlb = 0, sdrb = 0, isdb = 0, droeb = b[0];
    Bxib += b[5];
    // h
    const double Hl = Ul.e + Wl.p + pbl, Hr = Ur.e + Wr.p + pbr;
    const double Ql = Hl / sdl, Qr = Hr / sdr;
    const double Qb = b[4] * isd;
    isdb += b[4] * (Ql + Qr);
    const double Hlb = adiv(Qb, sdl), Hrb = adiv(Qb, sdr);
    sdlb -= adiv(Qb * Ql, sdl);
    sdrb -= adiv(Qb * Qr, sdr);
    Ulb.e += Hlb;
    Wlb.p += Hlb;
    pblb += Hlb;
    Urb.e += Hrb;
    Wrb.p += Hrb;
    pbrb += Hrb;
    Bxib += (pblb + pbrb) * Bxi;
    Wlb.by += pblb * Wl.by;
    Wlb.bz += pblb * Wl.bz;
    Wrb.by += pbrb * Wr.by;
    Wrb.bz += pbrb * Wr.bz;
    // y = 0.5 (dl + dr) / droe
    Wlb.d += adiv(0.5 * b[9], droe);
    Wrb.d += adiv(0.5 * b[9], droe);
    droeb -= adiv(b[9] * y, droe);
    // x = 0.5 (dby^2 + dbz^2) / S^2
    {
      const double S = sdl + sdr;
      const double Dsqb = adiv(b[8] * 0.5, S * S);
      const double Sb = -adiv(2.0 * b[8] * x, S);
      sdlb += Sb;
      sdrb += Sb;
      const double dby = Wl.by - Wr.by, dbz = Wl.bz - Wr.bz;
      Wlb.by += 2.0 * dby * Dsqb;
      Wrb.by -= 2.0 * dby * Dsqb;
      Wlb.bz += 2.0 * dbz * Dsqb;
      Wrb.bz -= 2.0 * dbz * Dsqb;
    }
    // b2, b3 (left value weighted by sqrt(dr))
    {
      const double N2b = b[6] * isd, N3b = b[7] * isd;
      isdb += b[6] * (sdr * Wl.by + sdl * Wr.by) + b[7] * (sdr * Wl.bz + sdl * Wr.bz);
      sdrb += N2b * Wl.by + N3b * Wl.bz;
      sdlb += N2b * Wr.by + N3b * Wr.bz;
      Wlb.by += N2b * sdr;
      Wrb.by += N2b * sdl;
      Wlb.bz += N3b * sdr;
      Wrb.bz += N3b * sdl;
    }
    // v1, v2, v3
    {
      const double n1 = b[1] * isd, n2 = b[2] * isd, n3 = b[3] * isd;
      isdb += b[1] * (sdl * Wl.vx + sdr * Wr.vx) + b[2] * (sdl * Wl.vy + sdr * Wr.vy) +
              b[3] * (sdl * Wl.vz + sdr * Wr.vz);
      sdlb += n1 * Wl.vx + n2 * Wl.vy + n3 * Wl.vz;
      sdrb += n1 * Wr.vx + n2 * Wr.vy + n3 * Wr.vz;
      Wlb.vx += n1 * sdl;
      Wlb.vy += n2 * sdl;
      Wlb.vz += n3 * sdl;
      Wrb.vx += n1 * sdr;
      Wrb.vy += n2 * sdr;
      Wrb.vz += n3 * sdr;
    }
    sdlb += droeb * sdr;
    sdrb += droeb * sdl;
    const double Tb = -isdb * isd * isd;
    sdlb += Tb;
    sdrb += Tb;
    Wlb.d += sqrt_b(sdl, sdlb);
    Wrb.d += sqrt_b(sdr, sdrb);
  }
};

// ---------------------------------------------------------------------------
// esystem_roe.c:56-292: forward record; VEC = false is the eigenvalues-only
// call (R == NULL, :145).
template <bool VEC>
struct EsysRoeF {
  double d, v1, v2, v3, h, b1, b2, b3, x, y, g1, g2;
  double di, vsq, btsq, gy, bt_starsq, vaxsq, hp, tas0, twid_asq, ct2, tsum, tdif, cf2_cs2;
  double cfsq, cf, cssq, cs, vax;
  double bt, bt_star, bet2, bet3, sgy, bet2_star, bet3_star, bet_starsq, vbet;
  double alpha_f, alpha_s, sqrtd, isqrtd, s, twid_a, qf, qs, af_prime, as_prime, afpbb, aspbb;
  double norm, cff, css, qfN, qsN, afN, asN, afpbL, aspbL, norm1, alf, als, q2s, q3s, vqstr;
  double norm2;
  int abr, btz;
  double ev[7];

  ATHMI_HDM void run(double d_, double v1_, double v2_, double v3_, double h_, double b1_,
                    double b2_, double b3_, double x_, double y_, const EosC& eos) {
    d = d_; v1 = v1_; v2 = v2_; v3 = v3_; h = h_; b1 = b1_; b2 = b2_; b3 = b3_; x = x_; y = y_;
    g1 = eos.g1;
    g2 = eos.g2;
    di = 1.0 / d;
    vsq = v1 * v1 + v2 * v2 + v3 * v3;
    btsq = b2 * b2 + b3 * b3;
    gy = g1 - g2 * y;
    bt_starsq = gy * btsq;
    vaxsq = b1 * b1 * di;
    hp = h - (vaxsq + btsq * di);
    tas0 = g1 * (hp - 0.5 * vsq) - g2 * x;
    twid_asq = dmax(tas0, kTiny);
    ct2 = bt_starsq * di;
    tsum = vaxsq + ct2 + twid_asq;
    tdif = vaxsq + ct2 - twid_asq;
    cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
    cfsq = 0.5 * (tsum + cf2_cs2);
    cf = sqrt(cfsq);
    cssq = twid_asq * vaxsq / cfsq;
    cs = sqrt(cssq);
    vax = sqrt(vaxsq);
    ev[0] = v1 - cf;
    ev[1] = v1 - vax;
    ev[2] = v1 - cs;
    ev[3] = v1;
    ev[4] = v1 + cs;
    ev[5] = v1 + vax;
    ev[6] = v1 + cf;
    if (!VEC) return;
    bt = sqrt(btsq);
    bt_star = sqrt(bt_starsq);
    btz = (bt == 0.0);
    if (btz) {
      bet2 = 1.0;
      bet3 = 0.0;
    } else {
      bet2 = b2 / bt;
      bet3 = b3 / bt;
    }
    sgy = sqrt(gy);
    bet2_star = bet2 / sgy;
    bet3_star = bet3 / sgy;
    bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
    vbet = v2 * bet2_star + v3 * bet3_star;
    abr = 0;
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
      abr = 1;
      alpha_f = sqrt((twid_asq - cssq) / (cfsq - cssq));
      alpha_s = sqrt((cfsq - twid_asq) / (cfsq - cssq));
    }
    sqrtd = sqrt(d);
    isqrtd = 1.0 / sqrtd;
    s = dsign(b1);
    twid_a = sqrt(twid_asq);
    qf = cf * alpha_f * s;
    qs = cs * alpha_s * s;
    af_prime = twid_a * alpha_f * isqrtd;
    as_prime = twid_a * alpha_s * isqrtd;
    afpbb = af_prime * bt_star * bet_starsq;
    aspbb = as_prime * bt_star * bet_starsq;
    norm = 0.5 / twid_asq;
    cff = norm * alpha_f * cf;
    css = norm * alpha_s * cs;
    qfN = qf * norm;
    qsN = qs * norm;
    afN = norm * af_prime * d;
    asN = norm * as_prime * d;
    afpbL = norm * af_prime * bt_star;
    aspbL = norm * as_prime * bt_star;
    norm1 = norm * g1;
    alf = alpha_f * norm1;
    als = alpha_s * norm1;
    q2s = bet2_star / bet_starsq;
    q3s = bet3_star / bet_starsq;
    vqstr = (v2 * q2s + v3 * q3s);
    norm2 = norm1 * 2.0;
  }

  ATHMI_HDM double R(int i, int j) const {
    switch (i * 7 + j) {
      case 0: case 6: return alpha_f;
      case 2: case 4: return alpha_s;
      case 3: return 1.0;
      case 7: return alpha_f * ev[0];
      case 9: return alpha_s * ev[2];
      case 10: return v1;
      case 11: return alpha_s * ev[4];
      case 13: return alpha_f * ev[6];
      case 14: return alpha_f * v2 + qs * bet2_star;
      case 15: return -bet3;
      case 16: return alpha_s * v2 - qf * bet2_star;
      case 17: return v2;
      case 18: return alpha_s * v2 + qf * bet2_star;
      case 19: return bet3;
      case 20: return alpha_f * v2 - qs * bet2_star;
      case 21: return alpha_f * v3 + qs * bet3_star;
      case 22: return bet2;
      case 23: return alpha_s * v3 - qf * bet3_star;
      case 24: return v3;
      case 25: return alpha_s * v3 + qf * bet3_star;
      case 26: return -bet2;
      case 27: return alpha_f * v3 - qs * bet3_star;
      case 28: return alpha_f * (hp - v1 * cf) + qs * vbet + aspbb;
      case 29: return -(v2 * bet3 - v3 * bet2);
      case 30: return alpha_s * (hp - v1 * cs) - qf * vbet - afpbb;
      case 31: return 0.5 * vsq + g2 * x / g1;
      case 32: return alpha_s * (hp + v1 * cs) + qf * vbet - afpbb;
      case 33: return (v2 * bet3 - v3 * bet2);
      case 34: return alpha_f * (hp + v1 * cf) - qs * vbet + aspbb;
      case 35: case 41: return as_prime * bet2_star;
      case 36: case 40: return -bet3 * s * isqrtd;
      case 37: case 39: return -af_prime * bet2_star;
      case 42: case 48: return as_prime * bet3_star;
      case 43: case 47: return bet2 * s * isqrtd;
      case 44: case 46: return -af_prime * bet3_star;
      default: return 0.0;
    }
  }
  ATHMI_HDM double L(int i, int j) const {
    switch (i * 7 + j) {
      case 0: return alf * (vsq - hp) + cff * (cf + v1) - qsN * vqstr - aspbL;
      case 1: return -alf * v1 - cff;
      case 2: return -alf * v2 + qsN * q2s;
      case 3: return -alf * v3 + qsN * q3s;
      case 4: case 46: return alf;
      case 5: case 47: return asN * q2s - alf * b2;
      case 6: case 48: return asN * q3s - alf * b3;
      case 7: return 0.5 * (v2 * bet3 - v3 * bet2);
      case 35: return -(0.5 * (v2 * bet3 - v3 * bet2));
      case 9: return -0.5 * bet3;
      case 37: return -(-0.5 * bet3);
      case 10: return 0.5 * bet2;
      case 38: return -(0.5 * bet2);
      case 12: case 40: return -0.5 * sqrtd * bet3 * s;
      case 13: case 41: return 0.5 * sqrtd * bet2 * s;
      case 14: return als * (vsq - hp) + css * (cs + v1) + qfN * vqstr + afpbL;
      case 15: return -als * v1 - css;
      case 16: return -als * v2 - qfN * q2s;
      case 17: return -als * v3 - qfN * q3s;
      case 18: case 32: return als;
      case 19: case 33: return -afN * q2s - als * b2;
      case 20: case 34: return -afN * q3s - als * b3;
      case 21: return 1.0 - norm2 * (0.5 * vsq - g2 * x / g1);
      case 22: return norm2 * v1;
      case 23: return norm2 * v2;
      case 24: return norm2 * v3;
      case 25: return -norm2;
      case 26: return norm2 * b2;
      case 27: return norm2 * b3;
      case 28: return als * (vsq - hp) + css * (cs - v1) - qfN * vqstr + afpbL;
      case 29: return -als * v1 + css;
      case 30: return -als * v2 + qfN * q2s;
      case 31: return -als * v3 + qfN * q3s;
      case 42: return alf * (vsq - hp) + cff * (cf - v1) + qsN * vqstr - aspbL;
      case 43: return -alf * v1 + cff;
      case 44: return -alf * v2 - qsN * q2s;
      case 45: return -alf * v3 - qsN * q3s;
      default: return 0.0;
    }
  }

  // Adjoint: Rb, Lb (ignored when !VEC), evb -> in_b[10] (+=) in the order
  // (d, v1, v2, v3, h, b1, b2, b3, x, y)
  template <class RB, class LB>
  ATHMI_HDM void back(const RB& Rb, const LB& Lb, const double* evb_in, double* in_b) const {
    double db = 0, v1b = 0, v2b = 0, v3b = 0, hb = 0, b1b = 0, b2b = 0, b3b = 0, xb = 0, yb = 0;
    double evb[7];
    for (int k = 0; k < 7; k++) evb[k] = evb_in[k];
    double hpb = 0, vsqb = 0, twid_asqb = 0, cssqb = 0, cfsqb = 0, cfb = 0, csb = 0, vaxb = 0;
    double dib = 0, btsqb = 0, bt_starsqb = 0, gyb = 0;
    if (VEC) {
      double alpha_fb = 0, alpha_sb = 0, bet2b = 0, bet3b = 0, qsb = 0, qfb = 0;
      double b2sb = 0, b3sb = 0, vbetb = 0, aspbbb = 0, afpbbb = 0, as_pb = 0, af_pb = 0;
      double isqrtdb = 0, sqrtdb = 0, alfb = 0, alsb = 0, cffb = 0, cssb = 0, qsNb = 0;
      double qfNb = 0, vqstrb = 0, aspbLb = 0, afpbLb = 0, q2sb = 0, q3sb = 0, asNb = 0;
      double afNb = 0, norm2b = 0, X;
      // ---- right eigenvectors
      alpha_fb += Rb(0, 0) + Rb(0, 6);
      alpha_sb += Rb(0, 2) + Rb(0, 4);
      alpha_fb += Rb(1, 0) * ev[0] + Rb(1, 6) * ev[6];
      evb[0] += Rb(1, 0) * alpha_f;
      evb[6] += Rb(1, 6) * alpha_f;
      alpha_sb += Rb(1, 2) * ev[2] + Rb(1, 4) * ev[4];
      evb[2] += Rb(1, 2) * alpha_s;
      evb[4] += Rb(1, 4) * alpha_s;
      v1b += Rb(1, 3);
      {
        const double qab = Rb(2, 0) + Rb(2, 6), qcb = Rb(2, 0) - Rb(2, 6);
        const double qbb = Rb(2, 2) + Rb(2, 4), qdb = Rb(2, 4) - Rb(2, 2);
        bet3b += Rb(2, 5) - Rb(2, 1);
        v2b += Rb(2, 3) + qab * alpha_f + qbb * alpha_s;
        alpha_fb += qab * v2;
        alpha_sb += qbb * v2;
        qsb += qcb * bet2_star;
        qfb += qdb * bet2_star;
        b2sb += qcb * qs + qdb * qf;
      }
      {
        const double qab = Rb(3, 0) + Rb(3, 6), qcb = Rb(3, 0) - Rb(3, 6);
        const double qbb = Rb(3, 2) + Rb(3, 4), qdb = Rb(3, 4) - Rb(3, 2);
        bet2b += Rb(3, 1) - Rb(3, 5);
        v3b += Rb(3, 3) + qab * alpha_f + qbb * alpha_s;
        alpha_fb += qab * v3;
        alpha_sb += qbb * v3;
        qsb += qcb * bet3_star;
        qfb += qdb * bet3_star;
        b3sb += qcb * qs + qdb * qf;
      }
      X = Rb(4, 5) - Rb(4, 1);  // (v2*bet3 - v3*bet2)
      v2b += X * bet3;
      bet3b += X * v2;
      v3b -= X * bet2;
      bet2b -= X * v3;
      vsqb += 0.5 * Rb(4, 3);
      xb += adiv(Rb(4, 3) * g2, g1);
      {
        const double r0 = Rb(4, 0), r6 = Rb(4, 6), r2 = Rb(4, 2), r4 = Rb(4, 4);
        alpha_fb += r0 * (hp - v1 * cf) + r6 * (hp + v1 * cf);
        alpha_sb += r2 * (hp - v1 * cs) + r4 * (hp + v1 * cs);
        hpb += (r0 + r6) * alpha_f + (r2 + r4) * alpha_s;
        v1b += (r6 - r0) * alpha_f * cf + (r4 - r2) * alpha_s * cs;
        cfb += (r6 - r0) * alpha_f * v1;
        csb += (r4 - r2) * alpha_s * v1;
        qsb += (r0 - r6) * vbet;
        qfb += (r4 - r2) * vbet;
        vbetb += (r0 - r6) * qs + (r4 - r2) * qf;
        aspbbb += r0 + r6;
        afpbbb -= r2 + r4;
      }
      X = Rb(5, 0) + Rb(5, 6);
      as_pb += X * bet2_star;
      b2sb += X * as_prime;
      X = Rb(5, 1) + Rb(5, 5);
      bet3b -= X * s * isqrtd;
      isqrtdb -= X * bet3 * s;
      X = Rb(5, 2) + Rb(5, 4);
      af_pb -= X * bet2_star;
      b2sb -= X * af_prime;
      X = Rb(6, 0) + Rb(6, 6);
      as_pb += X * bet3_star;
      b3sb += X * as_prime;
      X = Rb(6, 1) + Rb(6, 5);
      bet2b += X * s * isqrtd;
      isqrtdb += X * bet2 * s;
      X = Rb(6, 2) + Rb(6, 4);
      af_pb -= X * bet3_star;
      b3sb -= X * af_prime;
      // ---- left eigenvectors, rows 0 and 6
      {
        const double l0 = Lb(0, 0), l6 = Lb(6, 0);
        alfb += (l0 + l6) * (vsq - hp);
        vsqb += (l0 + l6) * alf;
        hpb -= (l0 + l6) * alf;
        cffb += l0 * (cf + v1) + l6 * (cf - v1);
        cfb += (l0 + l6) * cff;
        v1b += (l0 - l6) * cff;
        qsNb += (l6 - l0) * vqstr;
        vqstrb += (l6 - l0) * qsN;
        aspbLb -= l0 + l6;
      }
      X = Lb(0, 1) + Lb(6, 1);
      alfb -= X * v1;
      v1b -= X * alf;
      cffb += Lb(6, 1) - Lb(0, 1);
      X = Lb(0, 2) + Lb(6, 2);
      alfb -= X * v2;
      v2b -= X * alf;
      qsNb += (Lb(0, 2) - Lb(6, 2)) * q2s;
      q2sb += (Lb(0, 2) - Lb(6, 2)) * qsN;
      X = Lb(0, 3) + Lb(6, 3);
      alfb -= X * v3;
      v3b -= X * alf;
      qsNb += (Lb(0, 3) - Lb(6, 3)) * q3s;
      q3sb += (Lb(0, 3) - Lb(6, 3)) * qsN;
      alfb += Lb(0, 4) + Lb(6, 4);
      X = Lb(0, 5) + Lb(6, 5);
      asNb += X * q2s;
      q2sb += X * asN;
      alfb -= X * b2;
      b2b -= X * alf;
      X = Lb(0, 6) + Lb(6, 6);
      asNb += X * q3s;
      q3sb += X * asN;
      alfb -= X * b3;
      b3b -= X * alf;
      // rows 1 and 5
      X = 0.5 * (Lb(1, 0) - Lb(5, 0));
      v2b += X * bet3;
      bet3b += X * v2;
      v3b -= X * bet2;
      bet2b -= X * v3;
      bet3b -= 0.5 * (Lb(1, 2) - Lb(5, 2));
      bet2b += 0.5 * (Lb(1, 3) - Lb(5, 3));
      X = Lb(1, 5) + Lb(5, 5);
      sqrtdb -= 0.5 * X * bet3 * s;
      bet3b -= 0.5 * X * sqrtd * s;
      X = Lb(1, 6) + Lb(5, 6);
      sqrtdb += 0.5 * X * bet2 * s;
      bet2b += 0.5 * X * sqrtd * s;
      // rows 2 and 4
      {
        const double l2 = Lb(2, 0), l4 = Lb(4, 0);
        alsb += (l2 + l4) * (vsq - hp);
        vsqb += (l2 + l4) * als;
        hpb -= (l2 + l4) * als;
        cssb += l2 * (cs + v1) + l4 * (cs - v1);
        csb += (l2 + l4) * css;
        v1b += (l2 - l4) * css;
        qfNb += (l2 - l4) * vqstr;
        vqstrb += (l2 - l4) * qfN;
        afpbLb += l2 + l4;
      }
      X = Lb(2, 1) + Lb(4, 1);
      alsb -= X * v1;
      v1b -= X * als;
      cssb += Lb(4, 1) - Lb(2, 1);
      X = Lb(2, 2) + Lb(4, 2);
      alsb -= X * v2;
      v2b -= X * als;
      qfNb += (Lb(4, 2) - Lb(2, 2)) * q2s;
      q2sb += (Lb(4, 2) - Lb(2, 2)) * qfN;
      X = Lb(2, 3) + Lb(4, 3);
      alsb -= X * v3;
      v3b -= X * als;
      qfNb += (Lb(4, 3) - Lb(2, 3)) * q3s;
      q3sb += (Lb(4, 3) - Lb(2, 3)) * qfN;
      alsb += Lb(2, 4) + Lb(4, 4);
      X = Lb(2, 5) + Lb(4, 5);
      afNb -= X * q2s;
      q2sb -= X * afN;
      alsb -= X * b2;
      b2b -= X * als;
      X = Lb(2, 6) + Lb(4, 6);
      afNb -= X * q3s;
      q3sb -= X * afN;
      alsb -= X * b3;
      b3b -= X * als;
      // row 3
      norm2b -= Lb(3, 0) * (0.5 * vsq - adiv(g2 * x, g1));
      vsqb -= Lb(3, 0) * norm2 * 0.5;
      xb += adiv(Lb(3, 0) * norm2 * g2, g1);
      norm2b += Lb(3, 1) * v1 + Lb(3, 2) * v2 + Lb(3, 3) * v3 - Lb(3, 4) + Lb(3, 5) * b2 +
                Lb(3, 6) * b3;
      v1b += Lb(3, 1) * norm2;
      v2b += Lb(3, 2) * norm2;
      v3b += Lb(3, 3) * norm2;
      b2b += Lb(3, 5) * norm2;
      b3b += Lb(3, 6) * norm2;
      // ---- prefixes of L
      double norm1b = 2.0 * norm2b;
      v2b += vqstrb * q2s;
      q2sb += vqstrb * v2;
      v3b += vqstrb * q3s;
      q3sb += vqstrb * v3;
      double bet_starsqb = -adiv(q2sb * q2s + q3sb * q3s, bet_starsq);
      b2sb += adiv(q2sb, bet_starsq);
      b3sb += adiv(q3sb, bet_starsq);
      alpha_fb += alfb * norm1;
      alpha_sb += alsb * norm1;
      norm1b += alfb * alpha_f + alsb * alpha_s;
      double normb = norm1b * g1;
      normb += afpbLb * af_prime * bt_star + aspbLb * as_prime * bt_star;
      af_pb += afpbLb * norm * bt_star;
      as_pb += aspbLb * norm * bt_star;
      double bt_starb = afpbLb * norm * af_prime + aspbLb * norm * as_prime;
      normb += afNb * af_prime * d + asNb * as_prime * d;
      af_pb += afNb * norm * d;
      as_pb += asNb * norm * d;
      db += afNb * norm * af_prime + asNb * norm * as_prime;
      qfb += qfNb * norm;
      qsb += qsNb * norm;
      normb += qfNb * qf + qsNb * qs;
      normb += cffb * alpha_f * cf + cssb * alpha_s * cs;
      alpha_fb += cffb * norm * cf;
      cfb += cffb * norm * alpha_f;
      alpha_sb += cssb * norm * cs;
      csb += cssb * norm * alpha_s;
      twid_asqb -= adiv(normb * norm, twid_asq);
      // ---- afpbb, aspbb, af_prime, as_prime, qf, qs
      af_pb += afpbbb * bt_star * bet_starsq;
      as_pb += aspbbb * bt_star * bet_starsq;
      bt_starb += afpbbb * af_prime * bet_starsq + aspbbb * as_prime * bet_starsq;
      bet_starsqb += afpbbb * af_prime * bt_star + aspbbb * as_prime * bt_star;
      double twid_ab = af_pb * alpha_f * isqrtd + as_pb * alpha_s * isqrtd;
      alpha_fb += af_pb * twid_a * isqrtd;
      alpha_sb += as_pb * twid_a * isqrtd;
      isqrtdb += af_pb * twid_a * alpha_f + as_pb * twid_a * alpha_s;
      cfb += qfb * alpha_f * s;
      alpha_fb += qfb * cf * s;
      csb += qsb * alpha_s * s;
      alpha_sb += qsb * cs * s;
      twid_asqb += sqrt_b(twid_a, twid_ab);
      sqrtdb -= isqrtdb * isqrtd * isqrtd;
      db += sqrt_b(sqrtd, sqrtdb);
      if (abr) {
        const double den = cfsq - cssq;
        const double Qf = (twid_asq - cssq) / den, Qs = (cfsq - twid_asq) / den;
        const double Qfb = sqrt_b(alpha_f, alpha_fb), Qsb = sqrt_b(alpha_s, alpha_sb);
        twid_asqb += adiv(Qfb - Qsb, den);
        cssqb -= adiv(Qfb, den);
        cfsqb += adiv(Qsb, den);
        const double denb = -adiv(Qfb * Qf + Qsb * Qs, den);
        cfsqb += denb;
        cssqb -= denb;
      }
      v2b += vbetb * bet2_star;
      v3b += vbetb * bet3_star;
      b2sb += vbetb * v2 + 2.0 * bet2_star * bet_starsqb;
      b3sb += vbetb * v3 + 2.0 * bet3_star * bet_starsqb;
      bet2b += adiv(b2sb, sgy);
      bet3b += adiv(b3sb, sgy);
      const double sgyb = -adiv(b2sb * bet2_star + b3sb * bet3_star, sgy);
      gyb += sqrt_b(sgy, sgyb);
      if (!btz) {
        b2b += adiv(bet2b, bt);
        b3b += adiv(bet3b, bt);
        const double btb = -adiv(bet2b * bet2 + bet3b * bet3, bt);
        btsqb += sqrt_b(bt, btb);
      }
      bt_starsqb += sqrt_b(bt_star, bt_starb);
    }
    // ---- eigenvalues and the speeds
    for (int k = 0; k < 7; k++) v1b += evb[k];
    cfb += evb[6] - evb[0];
    vaxb += evb[5] - evb[1];
    csb += evb[4] - evb[2];
    double vaxsqb = sqrt_b(vax, vaxb);
    cssqb += sqrt_b(cs, csb);
    twid_asqb += adiv(cssqb * vaxsq, cfsq);
    vaxsqb += adiv(cssqb * twid_asq, cfsq);
    cfsqb -= adiv(cssqb * cssq, cfsq);
    cfsqb += sqrt_b(cf, cfb);
    const double tsumb = 0.5 * cfsqb;
    const double Yb = sqrt_b(cf2_cs2, 0.5 * cfsqb);
    const double tdifb = 2.0 * tdif * Yb;
    twid_asqb += 4.0 * ct2 * Yb;
    double ct2b = 4.0 * twid_asq * Yb;
    vaxsqb += tsumb + tdifb;
    ct2b += tsumb + tdifb;
    twid_asqb += tsumb - tdifb;
    bt_starsqb += ct2b * di;
    dib += ct2b * bt_starsq;
    const double tas0b = (tas0 > kTiny) ? twid_asqb : 0.0;
    hpb += g1 * tas0b;
    vsqb -= 0.5 * g1 * tas0b;
    xb -= g2 * tas0b;
    hb += hpb;
    vaxsqb -= hpb;
    btsqb -= hpb * di;
    dib -= hpb * btsq;
    b1b += vaxsqb * 2.0 * b1 * di;
    dib += vaxsqb * b1 * b1;
    gyb += bt_starsqb * btsq;
    btsqb += bt_starsqb * gy;
    yb -= g2 * gyb;
    b2b += 2.0 * b2 * btsqb;
    b3b += 2.0 * b3 * btsqb;
    v1b += 2.0 * v1 * vsqb;
    v2b += 2.0 * v2 * vsqb;
    v3b += 2.0 * v3 * vsqb;
    db -= dib * di * di;
    in_b[0] += db;
    in_b[1] += v1b;
    in_b[2] += v2b;
    in_b[3] += v3b;
    in_b[4] += hb;
    in_b[5] += b1b;
    in_b[6] += b2b;
    in_b[7] += b3b;
    in_b[8] += xb;
    in_b[9] += yb;
  }
};

struct ZeroMat {
  ATHMI_HDM double operator()(int, int) const { return 0.0; }
};

// ---------------------------------------------------------------------------
// flux_hlle.c:60-199 reversed
ATHMI_HD void hlle_speed_b(double p, double d, double Bxi, double uby, double ubz,
                           const EosC& eos, double cfb, double& pb, double& db, double& Bxib,
                           double& ubyb, double& ubzb) {
  const double asq = eos.gamma * p / d;
  const double vaxsq = Bxi * Bxi / d;
  const double ct2 = (uby * uby + ubz * ubz) / d;
  const double tmp = vaxsq + ct2 - asq;
  const double Z = sqrt(tmp * tmp + 4.0 * asq * ct2);
  const double cfsq = 0.5 * ((vaxsq + ct2 + asq) + Z);
  const double cfsqb = sqrt_b(sqrt(cfsq), cfb);
  const double qsqb = 0.5 * cfsqb;
  const double Yb = sqrt_b(Z, 0.5 * cfsqb);
  const double tmpb = 2.0 * tmp * Yb;
  const double asqb = 4.0 * ct2 * Yb + qsqb - tmpb;
  const double ct2b = 4.0 * asq * Yb + qsqb + tmpb;
  const double vaxsqb = qsqb + tmpb;
  pb += adiv(asqb * eos.gamma, d);
  db -= adiv(asqb * asq + vaxsqb * vaxsq + ct2b * ct2, d);
  Bxib += adiv(vaxsqb * 2.0 * Bxi, d);
  ubyb += adiv(ct2b * 2.0 * uby, d);
  ubzb += adiv(ct2b * 2.0 * ubz, d);
}

ATHMI_HD double hlle_speed(double p, double d, double Bxi, double uby, double ubz,
                           const EosC& eos) {
  double asq = eos.gamma * p / d;
  double vaxsq = Bxi * Bxi / d;
  double ct2 = (uby * uby + ubz * ubz) / d;
  double qsq = vaxsq + ct2 + asq;
  double tmp = vaxsq + ct2 - asq;
  double cfsq = 0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2));
  return sqrt(cfsq);
}

// HLLE side flux (flux_hlle.c:149-189) with speed b (bm on the left, bp on the
// right) and its adjoint
ATHMI_HD Cons7 hlle_side(const Cons7& U, const Prim7& W, double Bxi, double pb_, double b) {
  Cons7 F;
  F.d = U.mx - b * U.d;
  F.mx = U.mx * (W.vx - b);
  F.my = U.my * (W.vx - b);
  F.mz = U.mz * (W.vx - b);
  F.mx = F.mx + W.p;
  F.e = U.e * (W.vx - b) + W.p * W.vx;
  F.mx = F.mx - 0.5 * (Bxi * Bxi - sqr(W.by) - sqr(W.bz));
  F.my = F.my - Bxi * W.by;
  F.mz = F.mz - Bxi * W.bz;
  F.e = F.e + (pb_ * W.vx - Bxi * (Bxi * W.vx + W.by * W.vy + W.bz * W.vz));
  F.by = W.by * (W.vx - b) - Bxi * W.vy;
  F.bz = W.bz * (W.vx - b) - Bxi * W.vz;
  return F;
}
ATHMI_HD void hlle_side_b(const Cons7& U, const Prim7& W, double Bxi, double pb_, double b,
                          const Cons7& Fb, Cons7& Ub, Prim7& Wb, double& Bxib, double& pbb,
                          double& bb) {
  const double vl = W.vx - b;
  double vlb = 0.0;
  Ub.mx += Fb.d + Fb.mx * vl;
  bb -= Fb.d * U.d;
  Ub.d -= Fb.d * b;
  vlb += Fb.mx * U.mx + Fb.my * U.my + Fb.mz * U.mz + Fb.e * U.e + Fb.by * W.by + Fb.bz * W.bz;
  Wb.p += Fb.mx + Fb.e * W.vx;
  Bxib -= Fb.mx * Bxi;
  Wb.by += Fb.mx * W.by;
  Wb.bz += Fb.mx * W.bz;
  Ub.my += Fb.my * vl;
  Bxib -= Fb.my * W.by;
  Wb.by -= Fb.my * Bxi;
  Ub.mz += Fb.mz * vl;
  Bxib -= Fb.mz * W.bz;
  Wb.bz -= Fb.mz * Bxi;
  Ub.e += Fb.e * vl;
  Wb.vx += Fb.e * W.p + Fb.e * pb_;
  pbb += Fb.e * W.vx;
  const double K = Bxi * W.vx + W.by * W.vy + W.bz * W.vz;
  Bxib -= Fb.e * K;
  const double Kb = -Fb.e * Bxi;
  Bxib += Kb * W.vx;
  Wb.vx += Kb * Bxi;
  Wb.by += Kb * W.vy;
  Wb.vy += Kb * W.by;
  Wb.bz += Kb * W.vz;
  Wb.vz += Kb * W.bz;
  Wb.by += Fb.by * vl;
  Bxib -= Fb.by * W.vy;
  Wb.vy -= Fb.by * Bxi;
  Wb.bz += Fb.bz * vl;
  Bxib -= Fb.bz * W.vz;
  Wb.vz -= Fb.bz * Bxi;
  Wb.vx += vlb;
  bb -= vlb;
}

ATHMI_HD void flux_hlle_b(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                          double Bxi, const EosC& eos, const Cons7& Fb, Cons7& Ulb, Cons7& Urb,
                          Prim7& Wlb, Prim7& Wrb, double& Bxib) {
  RoeAvgF r;
  r.run(Ul, Ur, Wl, Wr, Bxi);
  EsysRoeF<false> e;
  e.run(r.droe, r.v1, r.v2, r.v3, r.h, Bxi, r.b2, r.b3, r.x, r.y, eos);
  const double cfl = hlle_speed(Wl.p, Wl.d, Bxi, Ul.by, Ul.bz, eos);
  const double cfr = hlle_speed(Wr.p, Wr.d, Bxi, Ur.by, Ur.bz, eos);
  const double ar = dmax(e.ev[6], Wr.vx + cfr);
  const double al = dmin(e.ev[0], Wl.vx - cfl);
  const double bp = dmax(ar, 0.0);
  const double bm = dmin(al, 0.0);
  const Cons7 Fl = hlle_side(Ul, Wl, Bxi, r.pbl, bm);
  const Cons7 Fr = hlle_side(Ur, Wr, Bxi, r.pbr, bp);
  const double tmp = 0.5 * (bp + bm) / (bp - bm);
  Cons7 Flb, Frb;
  double tmpb = 0.0;
  for (int n = 0; n < 7; n++) {
    arr(Flb)[n] = arr(Fb)[n] * (0.5 + tmp);
    arr(Frb)[n] = arr(Fb)[n] * (0.5 - tmp);
    tmpb += arr(Fb)[n] * (arr(Fl)[n] - arr(Fr)[n]);
  }
  const double den = (bp - bm) * (bp - bm);
  double bpb = -adiv(tmpb * bm, den), bmb = adiv(tmpb * bp, den);
  double pblb = 0.0, pbrb = 0.0;
  hlle_side_b(Ul, Wl, Bxi, r.pbl, bm, Flb, Ulb, Wlb, Bxib, pblb, bmb);
  hlle_side_b(Ur, Wr, Bxi, r.pbr, bp, Frb, Urb, Wrb, Bxib, pbrb, bpb);
  const double arb = (ar > 0.0) ? bpb : 0.0;
  const double alb = (al < 0.0) ? bmb : 0.0;
  double evb[7] = {0, 0, 0, 0, 0, 0, 0};
  double cfrb = 0.0, cflb = 0.0;
  if (e.ev[6] > Wr.vx + cfr) evb[6] += arb; else { Wrb.vx += arb; cfrb += arb; }
  if (e.ev[0] < Wl.vx - cfl) evb[0] += alb; else { Wlb.vx += alb; cflb -= alb; }
  hlle_speed_b(Wl.p, Wl.d, Bxi, Ul.by, Ul.bz, eos, cflb, Wlb.p, Wlb.d, Bxib, Ulb.by, Ulb.bz);
  hlle_speed_b(Wr.p, Wr.d, Bxi, Ur.by, Ur.bz, eos, cfrb, Wrb.p, Wrb.d, Bxib, Urb.by, Urb.bz);
  double inb[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  e.back(ZeroMat(), ZeroMat(), evb, inb);
  r.back(Ul, Ur, Wl, Wr, Bxi, inb, pblb, pbrb, Ulb, Urb, Wlb, Wrb, Bxib);
}

// ---------------------------------------------------------------------------
// flux_roe.c:54-259 reversed (etah = 0), with the HLLE fallback of :212-237
ATHMI_HD Cons7 roe_side(const Cons7& U, const Prim7& W, double Bxi, double pb_) {
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
  F.e = F.e + (pb_ * W.vx - Bxi * (Bxi * W.vx + W.by * W.vy + W.bz * W.vz));
  F.by = W.by * W.vx - Bxi * W.vy;
  F.bz = W.bz * W.vx - Bxi * W.vz;
  return F;
}
ATHMI_HD void roe_side_b(const Cons7& U, const Prim7& W, double Bxi, double pb_,
                         const Cons7& Fb, Cons7& Ub, Prim7& Wb, double& Bxib, double& pbb) {
  Ub.mx += Fb.d + Fb.mx * W.vx + Fb.my * W.vy + Fb.mz * W.vz;
  Wb.vx += Fb.mx * U.mx + Fb.e * (U.e + W.p) + Fb.e * pb_ + Fb.by * W.by + Fb.bz * W.bz;
  Wb.vy += Fb.my * U.mx;
  Wb.vz += Fb.mz * U.mx;
  Wb.p += Fb.mx + Fb.e * W.vx;
  Ub.e += Fb.e * W.vx;
  Bxib -= Fb.mx * Bxi + Fb.my * W.by + Fb.mz * W.bz;
  Wb.by += Fb.mx * W.by - Fb.my * Bxi;
  Wb.bz += Fb.mx * W.bz - Fb.mz * Bxi;
  pbb += Fb.e * W.vx;
  const double K = Bxi * W.vx + W.by * W.vy + W.bz * W.vz;
  Bxib -= Fb.e * K;
  const double Kb = -Fb.e * Bxi;
  Bxib += Kb * W.vx;
  Wb.vx += Kb * Bxi;
  Wb.by += Kb * W.vy;
  Wb.vy += Kb * W.by;
  Wb.bz += Kb * W.vz;
  Wb.vz += Kb * W.bz;
  Wb.by += Fb.by * W.vx;
  Wb.bz += Fb.bz * W.vx;
  Bxib -= Fb.by * W.vy + Fb.bz * W.vz;
  Wb.vy -= Fb.by * Bxi;
  Wb.vz -= Fb.bz * Bxi;
}

ATHMI_HD constexpr bool roe_Lnz(int n, int m) {
  return (n == 1 || n == 5) ? (m == 0 || m == 2 || m == 3 || m == 5 || m == 6) : true;
}
ATHMI_HD constexpr bool roe_Rnz(int n, int m) {
  return (n == 0) ? (m != 1 && m != 5)
         : (n == 1) ? (m != 1 && m != 5)
         : (n == 5 || n == 6) ? (m != 3)
                              : true;
}

// returns the fallback flag of the forward pass (0: Roe, 1/2: HLLE)
ATHMI_HD int flux_roe_b(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                        double Bxi, const EosC& eos, const Cons7& Fb, Cons7& Ulb, Cons7& Urb,
                        Prim7& Wlb, Prim7& Wrb, double& Bxib) {
  RoeAvgF r;
  r.run(Ul, Ur, Wl, Wr, Bxi);
  EsysRoeF<true> e;
  e.run(r.droe, r.v1, r.v2, r.v3, r.h, Bxi, r.b2, r.b3, r.x, r.y, eos);
  double pblb = 0.0, pbrb = 0.0;
  if (e.ev[0] >= 0.0) {  // :181
    roe_side_b(Ul, Wl, Bxi, r.pbl, Fb, Ulb, Wlb, Bxib, pblb);
    Bxib += pblb * Bxi;
    Wlb.by += pblb * Wl.by;
    Wlb.bz += pblb * Wl.bz;
    return 0;
  }
  if (e.ev[6] <= 0.0) {
    roe_side_b(Ur, Wr, Bxi, r.pbr, Fb, Urb, Wrb, Bxib, pbrb);
    Bxib += pbrb * Bxi;
    Wrb.by += pbrb * Wr.by;
    Wrb.bz += pbrb * Wr.bz;
    return 0;
  }
  double dU[7], a[7];
  for (int n = 0; n < 7; n++) dU[n] = arr(Ur)[n] - arr(Ul)[n];
  for (int n = 0; n < 7; n++) {
    double acc = 0.0;
    for (int m = 0; m < 7; m++)
      if (roe_Lnz(n, m)) acc = acc + e.L(n, m) * dU[m];
    a[n] = acc;
  }
  // intermediate-state positivity test (:212-237)
  int flag = 0;
  {
    double ui[7];
    for (int n = 0; n < 7; n++) ui[n] = arr(Ul)[n];
    for (int n = 0; n < 6 && !flag; n++) {
      for (int m = 0; m < 7; m++)
        if (roe_Rnz(m, n)) ui[m] = ui[m] + a[n] * e.R(m, n);
      if (e.ev[n + 1] > e.ev[n]) {
        if (ui[0] <= 0.0) {
          flag = 1;
          break;
        }
        double pi = ui[4] - 0.5 * (sqr(ui[1]) + sqr(ui[2]) + sqr(ui[3])) / ui[0];
        pi = pi - 0.5 * (sqr(ui[5]) + sqr(ui[6]) + sqr(Bxi));
        if (pi < 0.0) flag = 2;
      }
    }
  }
  if (flag) {
    flux_hlle_b(Ul, Ur, Wl, Wr, Bxi, eos, Fb, Ulb, Urb, Wlb, Wrb, Bxib);
    return flag;
  }
  double coeff[7], coeffb[7], ab[7], evb[7];
  for (int m = 0; m < 7; m++) coeff[m] = 0.5 * dmax(fabs(e.ev[m]), 0.0) * a[m];
  for (int m = 0; m < 7; m++) {
    double acc = 0.0;
    for (int n = 0; n < 7; n++)
      if (roe_Rnz(n, m)) acc -= arr(Fb)[n] * e.R(n, m);
    coeffb[m] = acc;
    const double mag = dmax(fabs(e.ev[m]), 0.0);
    ab[m] = coeffb[m] * 0.5 * mag;
    evb[m] = (fabs(e.ev[m]) > 0.0) ? coeffb[m] * 0.5 * a[m] * sgnabs(e.ev[m]) : 0.0;
  }
  for (int m = 0; m < 7; m++) {
    double acc = 0.0;
    for (int n = 0; n < 7; n++)
      if (roe_Lnz(n, m)) acc += ab[n] * e.L(n, m);
    arr(Urb)[m] += acc;
    arr(Ulb)[m] -= acc;
  }
  double inb[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  e.back([&](int n, int m) { return -arr(Fb)[n] * coeff[m]; },
         [&](int n, int m) { return ab[n] * dU[m]; }, evb, inb);
  Cons7 Fhb;
  for (int n = 0; n < 7; n++) arr(Fhb)[n] = 0.5 * arr(Fb)[n];
  roe_side_b(Ul, Wl, Bxi, r.pbl, Fhb, Ulb, Wlb, Bxib, pblb);
  roe_side_b(Ur, Wr, Bxi, r.pbr, Fhb, Urb, Wrb, Bxib, pbrb);
  r.back(Ul, Ur, Wl, Wr, Bxi, inb, pblb, pbrb, Ulb, Urb, Wlb, Wrb, Bxib);
  return 0;
}

// the solver fixed at compile time: a kernel then holds one reverse only (the
// Roe reverse keeps its HLLE fallback)
template <int KIND>
ATHMI_HD void riemann_b_of(const Cons7& Ul, const Cons7& Ur, const Prim7& Wl, const Prim7& Wr,
                           double Bxi, const EosC& eos, const Cons7& Fb, Cons7& Ulb, Cons7& Urb,
                           Prim7& Wlb, Prim7& Wrb, double& Bxib) {
  if constexpr (KIND == kFluxHlle)
    flux_hlle_b(Ul, Ur, Wl, Wr, Bxi, eos, Fb, Ulb, Urb, Wlb, Wrb, Bxib);
  else
    (void)flux_roe_b(Ul, Ur, Wl, Wr, Bxi, eos, Fb, Ulb, Urb, Wlb, Wrb, Bxib);
}

ATHMI_HD void riemann_b(int kind, const Cons7& Ul, const Cons7& Ur, const Prim7& Wl,
                        const Prim7& Wr, double Bxi, const EosC& eos, const Cons7& Fb,
                        Cons7& Ulb, Cons7& Urb, Prim7& Wlb, Prim7& Wrb, double& Bxib) {
  if (kind == kFluxHlle)
    flux_hlle_b(Ul, Ur, Wl, Wr, Bxi, eos, Fb, Ulb, Urb, Wlb, Wrb, Bxib);
  else
    (void)flux_roe_b(Ul, Ur, Wl, Wr, Bxi, eos, Fb, Ulb, Urb, Wlb, Wrb, Bxib);
}

}  // namespace adj
}  // namespace athmi
