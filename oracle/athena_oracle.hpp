// athena_oracle.hpp -- TEST INFRASTRUCTURE ONLY: the CPU restatement of the
// reference's 2-D CTU+CT step, used as the checker for the HIP path. Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Every routine restates one reference routine (file:line cited at each) with
// the same floating-point evaluation order, so that built with
// -ffp-contract=off it reproduces the reference bit for bit. It is templated on
// the scalar type: T=double is the oracle proper, T=Dual gives the
// tangent-linear model of the reference routines, T=Counted counts the FP64
// operations per cell-update.
//
// Parity pins (see tests/): every routine here is compared bit-exactly with the
// reference's own sources compiled into oracle/_ref, and with golden vectors
// written from them (tests/golden/).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace ora {

enum { NWAVE = 7, NGHOST = 4 };
static const double TINY_NUMBER = 1.0e-20;  // defs.h:103
static const double PI = 3.14159265358979323846;  // defs.h:99

// primal value of a scalar (identity for double; overloaded for AD types)
inline double pv(double x) { return x; }

// defs.h:89-96 macro semantics: MAX(a,b) = a>b ? a : b, MIN(a,b) = a<b ? a : b,
// SIGN(a) = a<0 ? -1 : +1 (so SIGN(-0.0) = +1)
template <class T> inline T MAXT(const T& a, const T& b) { return (pv(a) > pv(b)) ? a : b; }
template <class T> inline T MINT(const T& a, const T& b) { return (pv(a) < pv(b)) ? a : b; }
template <class T> inline double SIGNT(const T& a) { return (pv(a) < 0.) ? -1. : 1.; }
template <class T> inline T SQR(const T& a) { return a * a; }

struct Eos {
  double gamma, g1, g2;  // Gamma, Gamma_1, Gamma_2 (globals.h:18-31, main.c:92-94)
  explicit Eos(double g = 5.0 / 3.0) : gamma(g), g1(g - 1.0), g2(g - 2.0) {}
};

// athena.h:75-83 (Cons1D) and athena.h:89-97 (Prim1D): member order is the ABI
template <class T> struct Cons { T d, Mx, My, Mz, E, By, Bz; };
template <class T> struct Prim { T d, Vx, Vy, Vz, P, By, Bz; };
template <class T> inline T* as_arr(Cons<T>& c) { return &c.d; }
template <class T> inline const T* as_arr(const Cons<T>& c) { return &c.d; }
template <class T> inline T* as_arr(Prim<T>& c) { return &c.d; }
template <class T> inline const T* as_arr(const Prim<T>& c) { return &c.d; }

// ---------------------------------------------------------------------------
// convert_var.c:37-55
template <class T>
inline Prim<T> cons_to_prim(const Cons<T>& U, const T& Bx, const Eos& eos) {
  using std::sqrt;
  Prim<T> W;
  T di = 1.0 / U.d;
  W.d = U.d;
  W.Vx = U.Mx * di;
  W.Vy = U.My * di;
  W.Vz = U.Mz * di;
  W.P = U.E - 0.5 * (SQR(U.Mx) + SQR(U.My) + SQR(U.Mz)) * di;
  W.P = W.P - 0.5 * (SQR(Bx) + SQR(U.By) + SQR(U.Bz));
  W.P = W.P * eos.g1;
  W.P = MAXT(W.P, T(TINY_NUMBER));
  W.By = U.By;
  W.Bz = U.Bz;
  return W;
}

// convert_var.c:64-92
template <class T>
inline Cons<T> prim_to_cons(const Prim<T>& W, const T& Bx, const Eos& eos) {
  Cons<T> U;
  U.d = W.d;
  U.Mx = W.d * W.Vx;
  U.My = W.d * W.Vy;
  U.Mz = W.d * W.Vz;
  U.E = W.P / eos.g1 + 0.5 * W.d * (SQR(W.Vx) + SQR(W.Vy) + SQR(W.Vz));
  U.E = U.E + 0.5 * (SQR(Bx) + SQR(W.By) + SQR(W.Bz));
  U.By = W.By;
  U.Bz = W.Bz;
  return U;
}

// ---------------------------------------------------------------------------
// esystem_prim.c:36-235: eigenvalues and eigenvectors of the primitive-variable
// MHD Jacobian. R is column-wise (R[*][n] is the n-th right eigenvector), L is
// row-wise. Entries that are structurally zero are not written: callers pass
// zeroed matrices (lr_states_plm.c:83-88).
template <class T>
void esys_prim(const T& d, const T& v1, const T& p, const T& b1, const T& b2,
               const T& b3, const Eos& eos, T ev[7], T R[7][7], T L[7][7]) {
  using std::sqrt;
  T di, cfsq, cf, cssq, cs, bt, bet2, bet3, alpha_f, alpha_s;
  T sqrtd, a, qf, qs, af, as, vax, na, af_prime, as_prime;
  T tsum, tdif, cf2_cs2, ct2, btsq, vaxsq, asq;
  double s;
  di = 1.0 / d;
  btsq = b2 * b2 + b3 * b3;
  vaxsq = b1 * b1 * di;
  asq = eos.gamma * p * di;
  ct2 = btsq * di;                                     // :52
  tsum = vaxsq + ct2 + asq;
  tdif = vaxsq + ct2 - asq;
  cf2_cs2 = sqrt(tdif * tdif + 4.0 * asq * ct2);
  cfsq = 0.5 * (tsum + cf2_cs2);
  cf = sqrt(cfsq);
  cssq = asq * vaxsq / cfsq;
  cs = sqrt(cssq);
  bt = sqrt(btsq);                                     // :65
  if (pv(bt) == 0.0) {
    bet2 = T(1.0);
    bet3 = T(0.0);
  } else {
    bet2 = b2 / bt;
    bet3 = b3 / bt;
  }
  if (pv(cf2_cs2) == 0.0) {                            // :76
    alpha_f = T(1.0);
    alpha_s = T(0.0);
  } else if (pv(asq - cssq) <= 0.0) {
    alpha_f = T(0.0);
    alpha_s = T(1.0);
  } else if (pv(cfsq - asq) <= 0.0) {
    alpha_f = T(1.0);
    alpha_s = T(0.0);
  } else {
    alpha_f = sqrt((asq - cssq) / cf2_cs2);
    alpha_s = sqrt((cfsq - asq) / cf2_cs2);
  }
  sqrtd = sqrt(d);                                     // :92
  s = SIGNT(b1);
  a = sqrt(asq);
  qf = cf * alpha_f * s;
  qs = cs * alpha_s * s;
  af = a * alpha_f * sqrtd;
  as = a * alpha_s * sqrtd;
  vax = sqrt(vaxsq);                                   // :102
  ev[0] = v1 - cf;
  ev[1] = v1 - vax;
  ev[2] = v1 - cs;
  ev[3] = v1;
  ev[4] = v1 + cs;
  ev[5] = v1 + vax;
  ev[6] = v1 + cf;
  // right eigenvectors (:116-170)
  R[0][0] = d * alpha_f;
  R[0][2] = d * alpha_s;
  R[0][3] = T(1.0);
  R[0][4] = R[0][2];
  R[0][6] = R[0][0];
  R[1][0] = -cf * alpha_f;
  R[1][2] = -cs * alpha_s;
  R[1][4] = -R[1][2];
  R[1][6] = -R[1][0];
  R[2][0] = qs * bet2;
  R[2][1] = -bet3;
  R[2][2] = -qf * bet2;
  R[2][4] = -R[2][2];
  R[2][5] = bet3;
  R[2][6] = -R[2][0];
  R[3][0] = qs * bet3;
  R[3][1] = bet2;
  R[3][2] = -qf * bet3;
  R[3][4] = -R[3][2];
  R[3][5] = -bet2;
  R[3][6] = -R[3][0];
  R[4][0] = d * asq * alpha_f;
  R[4][2] = d * asq * alpha_s;
  R[4][4] = R[4][2];
  R[4][6] = R[4][0];
  R[5][0] = as * bet2;
  R[5][1] = -bet3 * s * sqrtd;
  R[5][2] = -af * bet2;
  R[5][4] = R[5][2];
  R[5][5] = R[5][1];
  R[5][6] = R[5][0];
  R[6][0] = as * bet3;
  R[6][1] = bet2 * s * sqrtd;
  R[6][2] = -af * bet3;
  R[6][4] = R[6][2];
  R[6][5] = R[6][1];
  R[6][6] = R[6][0];
  // left eigenvectors (:174-234)
  na = 0.5 / asq;
  qf = na * qf;
  qs = na * qs;
  af_prime = na * af * di;
  as_prime = na * as * di;
  L[0][1] = -na * cf * alpha_f;
  L[0][2] = qs * bet2;
  L[0][3] = qs * bet3;
  L[0][4] = na * alpha_f * di;
  L[0][5] = as_prime * bet2;
  L[0][6] = as_prime * bet3;
  L[1][2] = -0.5 * bet3;
  L[1][3] = 0.5 * bet2;
  L[1][5] = -0.5 * bet3 * s / sqrtd;
  L[1][6] = 0.5 * bet2 * s / sqrtd;
  L[2][1] = -na * cs * alpha_s;
  L[2][2] = -qf * bet2;
  L[2][3] = -qf * bet3;
  L[2][4] = na * alpha_s * di;
  L[2][5] = -af_prime * bet2;
  L[2][6] = -af_prime * bet3;
  L[3][0] = T(1.0);
  L[3][4] = -1.0 / asq;
  L[4][1] = -L[2][1];
  L[4][2] = -L[2][2];
  L[4][3] = -L[2][3];
  L[4][4] = L[2][4];
  L[4][5] = L[2][5];
  L[4][6] = L[2][6];
  L[5][2] = -L[1][2];
  L[5][3] = -L[1][3];
  L[5][5] = L[1][5];
  L[5][6] = L[1][6];
  L[6][1] = -L[0][1];
  L[6][2] = -L[0][2];
  L[6][3] = -L[0][3];
  L[6][4] = L[0][4];
  L[6][5] = L[0][5];
  L[6][6] = L[0][6];
}

// ---------------------------------------------------------------------------
// lr_states_plm.c:64-248. W and Bxc are indexed like the reference slice
// (valid on [il-2, iu+2]); writes Wl[il..iu+2], Wr[il-1..iu+1] (the reference's
// loop runs i = il-1 .. iu+1 and stores Wl[i+1], Wr[i]).
template <class T>
void lr_states(const Prim<T>* W, const T* Bxc, double dt, double dtodx, int il,
               int iu, Prim<T>* Wl, Prim<T>* Wr, const Eos& eos) {
  using std::fabs;
  (void)dt;
  T ev[NWAVE], rem[NWAVE][NWAVE], lem[NWAVE][NWAVE];
  T dWc[NWAVE], dWl[NWAVE], dWr[NWAVE], dWg[NWAVE];
  T dac[NWAVE], dal[NWAVE], dar[NWAVE], dag[NWAVE], da[NWAVE];
  T Wlv[NWAVE], Wrv[NWAVE], dW[NWAVE], dWm[NWAVE];
  T lim_slope1, lim_slope2, qa, qx;
  for (int n = 0; n < NWAVE; n++)
    for (int m = 0; m < NWAVE; m++) {
      rem[n][m] = T(0.0);
      lem[n][m] = T(0.0);
    }
  for (int i = il - 1; i <= iu + 1; i++) {
    const T* wm = as_arr(W[i - 1]);
    const T* w0 = as_arr(W[i]);
    const T* wp = as_arr(W[i + 1]);
    esys_prim(W[i].d, W[i].Vx, W[i].P, Bxc[i], W[i].By, W[i].Bz, eos, ev, rem, lem);
    for (int n = 0; n < NWAVE; n++) {                       // Step 2 (:108-117)
      dWc[n] = wp[n] - wm[n];
      dWl[n] = w0[n] - wm[n];
      dWr[n] = wp[n] - w0[n];
      if (pv(dWl[n] * dWr[n]) > 0.0)
        dWg[n] = 2.0 * dWl[n] * dWr[n] / (dWl[n] + dWr[n]);
      else
        dWg[n] = T(0.0);
    }
    for (int n = 0; n < NWAVE; n++) {                       // Step 3 (:122-133)
      dac[n] = lem[n][0] * dWc[0];
      dal[n] = lem[n][0] * dWl[0];
      dar[n] = lem[n][0] * dWr[0];
      dag[n] = lem[n][0] * dWg[0];
      for (int m = 1; m < NWAVE; m++) {
        dac[n] = dac[n] + lem[n][m] * dWc[m];
        dal[n] = dal[n] + lem[n][m] * dWl[m];
        dar[n] = dar[n] + lem[n][m] * dWr[m];
        dag[n] = dag[n] + lem[n][m] * dWg[m];
      }
    }
    for (int n = 0; n < NWAVE; n++) {                       // Step 4 (:138-145)
      da[n] = T(0.0);
      if (pv(dal[n] * dar[n]) > 0.0) {
        lim_slope1 = MINT(T(fabs(dal[n])), T(fabs(dar[n])));
        lim_slope2 = MINT(T(0.5 * fabs(dac[n])), T(fabs(dag[n])));
        da[n] = SIGNT(dac[n]) * MINT(T(2.0 * lim_slope1), lim_slope2);
      }
    }
    for (int n = 0; n < NWAVE; n++) {                       // Step 5 (:150-155)
      dWm[n] = da[0] * rem[n][0];
      for (int m = 1; m < NWAVE; m++) dWm[n] = dWm[n] + da[m] * rem[n][m];
    }
    for (int n = 0; n < NWAVE; n++) {                       // Step 7 (:160-174)
      Wlv[n] = w0[n] - 0.5 * dWm[n];
      Wrv[n] = w0[n] + 0.5 * dWm[n];
    }
    for (int n = 0; n < NWAVE; n++) {
      Wlv[n] = MAXT(MINT(w0[n], wm[n]), Wlv[n]);
      Wlv[n] = MINT(MAXT(w0[n], wm[n]), Wlv[n]);
      Wrv[n] = MAXT(MINT(w0[n], wp[n]), Wrv[n]);
      Wrv[n] = MINT(MAXT(w0[n], wp[n]), Wrv[n]);
    }
    for (int n = 0; n < NWAVE; n++) dW[n] = Wrv[n] - Wlv[n];
    T* pWl = as_arr(Wl[i + 1]);                              // Step 8 (:181-192)
    T* pWr = as_arr(Wr[i]);
    qx = 0.5 * MAXT(ev[NWAVE - 1], T(0.0)) * dtodx;
    for (int n = 0; n < NWAVE; n++) pWl[n] = Wrv[n] - qx * dW[n];
    qx = -0.5 * MINT(ev[0], T(0.0)) * dtodx;
    for (int n = 0; n < NWAVE; n++) pWr[n] = Wlv[n] + qx * dW[n];
    for (int n = 0; n < NWAVE; n++) {                       // Step 9 (:201-209)
      if (pv(ev[n]) > 0.) {
        qa = T(0.0);
        for (int m = 0; m < NWAVE; m++)
          qa = qa + lem[n][m] * 0.5 * dtodx * (ev[NWAVE - 1] - ev[n]) * dW[m];
        for (int m = 0; m < NWAVE; m++) pWl[m] = pWl[m] + qa * rem[m][n];
      }
    }
    for (int n = 0; n < NWAVE; n++) {                       // (:211-224)
      if (pv(ev[n]) < 0.) {
        qa = T(0.0);
        for (int m = 0; m < NWAVE; m++)
          qa = qa + lem[n][m] * 0.5 * dtodx * (ev[0] - ev[n]) * dW[m];
        for (int m = 0; m < NWAVE; m++) pWr[m] = pWr[m] + qa * rem[m][n];
        qa = T(0.0);
        for (int m = 0; m < NWAVE; m++)
          qa = qa + lem[n][m] * 0.5 * dtodx * (ev[n] - ev[NWAVE - 1]) * dW[m];
        for (int m = 0; m < NWAVE; m++) pWl[m] = pWl[m] - qa * rem[m][n];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// esystem_roe.c:56-292: Cargo-Gallice Roe-matrix eigensystem. With R == nullptr
// only the eigenvalues are computed (:145).
template <class T>
void esys_roe(const T& d, const T& v1, const T& v2, const T& v3, const T& h,
              const T& b1, const T& b2, const T& b3, const T& x, const T& y,
              const Eos& eos, T ev[7], T (*R)[7], T (*L)[7]) {
  using std::sqrt;
  const double Gamma_1 = eos.g1, Gamma_2 = eos.g2;
  T di, vsq, btsq, bt_starsq, vaxsq, hp, twid_asq, cfsq, cf, cssq, cs;
  T bt, bt_star, bet2, bet3, bet2_star, bet3_star, bet_starsq, vbet, alpha_f, alpha_s;
  T isqrtd, sqrtd, twid_a, qf, qs, af_prime, as_prime, afpbb, aspbb, vax;
  T norm, cff, css, af, as, afpb, aspb, q2_star, q3_star, vqstr;
  T ct2, tsum, tdif, cf2_cs2, qa, qb, qc, qd;
  double s;
  di = 1.0 / d;
  vsq = v1 * v1 + v2 * v2 + v3 * v3;
  btsq = b2 * b2 + b3 * b3;
  bt_starsq = (Gamma_1 - Gamma_2 * y) * btsq;
  vaxsq = b1 * b1 * di;
  hp = h - (vaxsq + btsq * di);
  twid_asq = MAXT(T(Gamma_1 * (hp - 0.5 * vsq) - Gamma_2 * x), T(TINY_NUMBER));  // :74
  ct2 = bt_starsq * di;
  tsum = vaxsq + ct2 + twid_asq;
  tdif = vaxsq + ct2 - twid_asq;
  cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
  cfsq = 0.5 * (tsum + cf2_cs2);
  cf = sqrt(cfsq);
  cssq = twid_asq * vaxsq / cfsq;
  cs = sqrt(cssq);
  bt = sqrt(btsq);                                              // :91
  bt_star = sqrt(bt_starsq);
  if (pv(bt) == 0.0) {
    bet2 = T(1.0);
    bet3 = T(0.0);
  } else {
    bet2 = b2 / bt;
    bet3 = b3 / bt;
  }
  bet2_star = bet2 / sqrt(Gamma_1 - Gamma_2 * y);
  bet3_star = bet3 / sqrt(Gamma_1 - Gamma_2 * y);
  bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
  vbet = v2 * bet2_star + v3 * bet3_star;
  if (pv(cfsq - cssq) == 0.0) {                                 // :107
    alpha_f = T(1.0);
    alpha_s = T(0.0);
  } else if (pv(twid_asq - cssq) <= 0.0) {
    alpha_f = T(0.0);
    alpha_s = T(1.0);
  } else if (pv(cfsq - twid_asq) <= 0.0) {
    alpha_f = T(1.0);
    alpha_s = T(0.0);
  } else {
    alpha_f = sqrt((twid_asq - cssq) / (cfsq - cssq));
    alpha_s = sqrt((cfsq - twid_asq) / (cfsq - cssq));
  }
  sqrtd = sqrt(d);                                              // :123
  isqrtd = 1.0 / sqrtd;
  s = SIGNT(b1);
  twid_a = sqrt(twid_asq);
  qf = cf * alpha_f * s;
  qs = cs * alpha_s * s;
  af_prime = twid_a * alpha_f * isqrtd;
  as_prime = twid_a * alpha_s * isqrtd;
  afpbb = af_prime * bt_star * bet_starsq;
  aspbb = as_prime * bt_star * bet_starsq;
  vax = sqrt(vaxsq);                                            // :136
  ev[0] = v1 - cf;
  ev[1] = v1 - vax;
  ev[2] = v1 - cs;
  ev[3] = v1;
  ev[4] = v1 + cs;
  ev[5] = v1 + vax;
  ev[6] = v1 + cf;
  if (R == nullptr || L == nullptr) return;
  R[0][0] = alpha_f;                                            // :151
  R[0][2] = alpha_s;
  R[0][3] = T(1.0);
  R[0][4] = alpha_s;
  R[0][6] = alpha_f;
  R[1][0] = alpha_f * ev[0];
  R[1][2] = alpha_s * ev[2];
  R[1][3] = v1;
  R[1][4] = alpha_s * ev[4];
  R[1][6] = alpha_f * ev[6];
  qa = alpha_f * v2;
  qb = alpha_s * v2;
  qc = qs * bet2_star;
  qd = qf * bet2_star;
  R[2][0] = qa + qc;
  R[2][1] = -bet3;
  R[2][2] = qb - qd;
  R[2][3] = v2;
  R[2][4] = qb + qd;
  R[2][5] = bet3;
  R[2][6] = qa - qc;
  qa = alpha_f * v3;
  qb = alpha_s * v3;
  qc = qs * bet3_star;
  qd = qf * bet3_star;
  R[3][0] = qa + qc;
  R[3][1] = bet2;
  R[3][2] = qb - qd;
  R[3][3] = v3;
  R[3][4] = qb + qd;
  R[3][5] = -bet2;
  R[3][6] = qa - qc;
  R[4][0] = alpha_f * (hp - v1 * cf) + qs * vbet + aspbb;       // :191
  R[4][1] = -(v2 * bet3 - v3 * bet2);
  R[4][2] = alpha_s * (hp - v1 * cs) - qf * vbet - afpbb;
  R[4][3] = 0.5 * vsq + Gamma_2 * x / Gamma_1;
  R[4][4] = alpha_s * (hp + v1 * cs) + qf * vbet - afpbb;
  R[4][5] = -R[4][1];
  R[4][6] = alpha_f * (hp + v1 * cf) - qs * vbet + aspbb;
  R[5][0] = as_prime * bet2_star;
  R[5][1] = -bet3 * s * isqrtd;
  R[5][2] = -af_prime * bet2_star;
  R[5][4] = R[5][2];
  R[5][5] = R[5][1];
  R[5][6] = R[5][0];
  R[6][0] = as_prime * bet3_star;
  R[6][1] = bet2 * s * isqrtd;
  R[6][2] = -af_prime * bet3_star;
  R[6][4] = R[6][2];
  R[6][5] = R[6][1];
  R[6][6] = R[6][0];
  norm = 0.5 / twid_asq;                                        // :218
  cff = norm * alpha_f * cf;
  css = norm * alpha_s * cs;
  qf = qf * norm;
  qs = qs * norm;
  af = norm * af_prime * d;
  as = norm * as_prime * d;
  afpb = norm * af_prime * bt_star;
  aspb = norm * as_prime * bt_star;
  norm = norm * Gamma_1;
  alpha_f = alpha_f * norm;
  alpha_s = alpha_s * norm;
  q2_star = bet2_star / bet_starsq;
  q3_star = bet3_star / bet_starsq;
  vqstr = (v2 * q2_star + v3 * q3_star);
  norm = norm * 2.0;
  L[0][0] = alpha_f * (vsq - hp) + cff * (cf + v1) - qs * vqstr - aspb;  // :237
  L[0][1] = -alpha_f * v1 - cff;
  L[0][2] = -alpha_f * v2 + qs * q2_star;
  L[0][3] = -alpha_f * v3 + qs * q3_star;
  L[0][4] = alpha_f;
  L[0][5] = as * q2_star - alpha_f * b2;
  L[0][6] = as * q3_star - alpha_f * b3;
  L[1][0] = 0.5 * (v2 * bet3 - v3 * bet2);
  L[1][2] = -0.5 * bet3;
  L[1][3] = 0.5 * bet2;
  L[1][5] = -0.5 * sqrtd * bet3 * s;
  L[1][6] = 0.5 * sqrtd * bet2 * s;
  L[2][0] = alpha_s * (vsq - hp) + css * (cs + v1) + qf * vqstr + afpb;
  L[2][1] = -alpha_s * v1 - css;
  L[2][2] = -alpha_s * v2 - qf * q2_star;
  L[2][3] = -alpha_s * v3 - qf * q3_star;
  L[2][4] = alpha_s;
  L[2][5] = -af * q2_star - alpha_s * b2;
  L[2][6] = -af * q3_star - alpha_s * b3;
  L[3][0] = 1.0 - norm * (0.5 * vsq - Gamma_2 * x / Gamma_1);
  L[3][1] = norm * v1;
  L[3][2] = norm * v2;
  L[3][3] = norm * v3;
  L[3][4] = -norm;
  L[3][5] = norm * b2;
  L[3][6] = norm * b3;
  L[4][0] = alpha_s * (vsq - hp) + css * (cs - v1) - qf * vqstr + afpb;
  L[4][1] = -alpha_s * v1 + css;
  L[4][2] = -alpha_s * v2 + qf * q2_star;
  L[4][3] = -alpha_s * v3 + qf * q3_star;
  L[4][4] = alpha_s;
  L[4][5] = L[2][5];
  L[4][6] = L[2][6];
  L[5][0] = -L[1][0];
  L[5][2] = -L[1][2];
  L[5][3] = -L[1][3];
  L[5][5] = L[1][5];
  L[5][6] = L[1][6];
  L[6][0] = alpha_f * (vsq - hp) + cff * (cf - v1) + qs * vqstr - aspb;
  L[6][1] = -alpha_f * v1 + cff;
  L[6][2] = -alpha_f * v2 - qs * q2_star;
  L[6][3] = -alpha_f * v3 - qs * q3_star;
  L[6][4] = alpha_f;
  L[6][5] = L[0][5];
  L[6][6] = L[0][6];
}

// ---------------------------------------------------------------------------
// Roe averages shared by both solvers (flux_roe.c:91-114, flux_hlle.c:87-110)
template <class T> struct RoeAvg {
  T sqrtdl, sqrtdr, isdlpdr, droe, v1roe, v2roe, v3roe, b2roe, b3roe, x, y, pbl, pbr, hroe;
};
template <class T>
inline RoeAvg<T> roe_average(const Cons<T>& Ul, const Cons<T>& Ur, const Prim<T>& Wl,
                             const Prim<T>& Wr, const T& Bxi) {
  using std::sqrt;
  RoeAvg<T> r;
  r.sqrtdl = sqrt(Wl.d);
  r.sqrtdr = sqrt(Wr.d);
  r.isdlpdr = 1.0 / (r.sqrtdl + r.sqrtdr);
  r.droe = r.sqrtdl * r.sqrtdr;
  r.v1roe = (r.sqrtdl * Wl.Vx + r.sqrtdr * Wr.Vx) * r.isdlpdr;
  r.v2roe = (r.sqrtdl * Wl.Vy + r.sqrtdr * Wr.Vy) * r.isdlpdr;
  r.v3roe = (r.sqrtdl * Wl.Vz + r.sqrtdr * Wr.Vz) * r.isdlpdr;
  r.b2roe = (r.sqrtdr * Wl.By + r.sqrtdl * Wr.By) * r.isdlpdr;
  r.b3roe = (r.sqrtdr * Wl.Bz + r.sqrtdl * Wr.Bz) * r.isdlpdr;
  r.x = 0.5 * (SQR(Wl.By - Wr.By) + SQR(Wl.Bz - Wr.Bz)) / (SQR(r.sqrtdl + r.sqrtdr));
  r.y = 0.5 * (Wl.d + Wr.d) / r.droe;
  r.pbl = 0.5 * (SQR(Bxi) + SQR(Wl.By) + SQR(Wl.Bz));
  r.pbr = 0.5 * (SQR(Bxi) + SQR(Wr.By) + SQR(Wr.Bz));
  r.hroe = ((Ul.E + Wl.P + r.pbl) / r.sqrtdl + (Ur.E + Wr.P + r.pbr) / r.sqrtdr) * r.isdlpdr;
  return r;
}

// flux_hlle.c:60-199
template <class T>
Cons<T> flux_hlle(const Cons<T>& Ul, const Cons<T>& Ur, const Prim<T>& Wl,
                  const Prim<T>& Wr, const T& Bxi, const Eos& eos) {
  using std::sqrt;
  RoeAvg<T> r = roe_average(Ul, Ur, Wl, Wr, Bxi);
  T ev[NWAVE];
  esys_roe<T>(r.droe, r.v1roe, r.v2roe, r.v3roe, r.hroe, Bxi, r.b2roe, r.b3roe, r.x, r.y,
              eos, ev, nullptr, nullptr);
  T asq, vaxsq, ct2, qsq, tmp, cfsq, cfl, cfr, ar, al, bp, bm;
  asq = eos.gamma * Wl.P / Wl.d;                                // :122
  vaxsq = Bxi * Bxi / Wl.d;
  ct2 = (Ul.By * Ul.By + Ul.Bz * Ul.Bz) / Wl.d;
  qsq = vaxsq + ct2 + asq;
  tmp = vaxsq + ct2 - asq;
  cfsq = 0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2));
  cfl = sqrt(cfsq);
  asq = eos.gamma * Wr.P / Wr.d;
  vaxsq = Bxi * Bxi / Wr.d;
  ct2 = (Ur.By * Ur.By + Ur.Bz * Ur.Bz) / Wr.d;
  qsq = vaxsq + ct2 + asq;
  tmp = vaxsq + ct2 - asq;
  cfsq = 0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2));
  cfr = sqrt(cfsq);
  ar = MAXT(ev[NWAVE - 1], T(Wr.Vx + cfr));                     // :139
  al = MINT(ev[0], T(Wl.Vx - cfl));
  bp = MAXT(ar, T(0.0));
  bm = MINT(al, T(0.0));
  Cons<T> Fl, Fr;
  Fl.d = Ul.Mx - bm * Ul.d;                                     // :149
  Fr.d = Ur.Mx - bp * Ur.d;
  Fl.Mx = Ul.Mx * (Wl.Vx - bm);
  Fr.Mx = Ur.Mx * (Wr.Vx - bp);
  Fl.My = Ul.My * (Wl.Vx - bm);
  Fr.My = Ur.My * (Wr.Vx - bp);
  Fl.Mz = Ul.Mz * (Wl.Vx - bm);
  Fr.Mz = Ur.Mz * (Wr.Vx - bp);
  Fl.Mx = Fl.Mx + Wl.P;
  Fr.Mx = Fr.Mx + Wr.P;
  Fl.E = Ul.E * (Wl.Vx - bm) + Wl.P * Wl.Vx;
  Fr.E = Ur.E * (Wr.Vx - bp) + Wr.P * Wr.Vx;
  Fl.Mx = Fl.Mx - 0.5 * (Bxi * Bxi - SQR(Wl.By) - SQR(Wl.Bz));
  Fr.Mx = Fr.Mx - 0.5 * (Bxi * Bxi - SQR(Wr.By) - SQR(Wr.Bz));
  Fl.My = Fl.My - Bxi * Wl.By;
  Fr.My = Fr.My - Bxi * Wr.By;
  Fl.Mz = Fl.Mz - Bxi * Wl.Bz;
  Fr.Mz = Fr.Mz - Bxi * Wr.Bz;
  Fl.E = Fl.E + (r.pbl * Wl.Vx - Bxi * (Bxi * Wl.Vx + Wl.By * Wl.Vy + Wl.Bz * Wl.Vz));
  Fr.E = Fr.E + (r.pbr * Wr.Vx - Bxi * (Bxi * Wr.Vx + Wr.By * Wr.Vy + Wr.Bz * Wr.Vz));
  Fl.By = Wl.By * (Wl.Vx - bm) - Bxi * Wl.Vy;
  Fr.By = Wr.By * (Wr.Vx - bp) - Bxi * Wr.Vy;
  Fl.Bz = Wl.Bz * (Wl.Vx - bm) - Bxi * Wl.Vz;
  Fr.Bz = Wr.Bz * (Wr.Vx - bp) - Bxi * Wr.Vz;
  Cons<T> F;
  const T* pFl = as_arr(Fl);
  const T* pFr = as_arr(Fr);
  T* pF = as_arr(F);
  tmp = 0.5 * (bp + bm) / (bp - bm);                            // :193
  for (int n = 0; n < NWAVE; n++) pF[n] = 0.5 * (pFl[n] + pFr[n]) + (pFl[n] - pFr[n]) * tmp;
  return F;
}

// flux_roe.c:54-259 (etah = 0: NO_H_CORRECTION, flux_roe.c:40)
template <class T>
Cons<T> flux_roe(const Cons<T>& Ul, const Cons<T>& Ur, const Prim<T>& Wl,
                 const Prim<T>& Wr, const T& Bxi, const Eos& eos, int* used_hlle = nullptr) {
  using std::fabs;
  const double etah = 0.0;
  T ev[NWAVE], rem[NWAVE][NWAVE], lem[NWAVE][NWAVE];
  for (int n = 0; n < NWAVE; n++)
    for (int m = 0; m < NWAVE; m++) {
      rem[n][m] = T(0.0);
      lem[n][m] = T(0.0);
    }
  RoeAvg<T> r = roe_average(Ul, Ur, Wl, Wr, Bxi);
  esys_roe<T>(r.droe, r.v1roe, r.v2roe, r.v3roe, r.hroe, Bxi, r.b2roe, r.b3roe, r.x, r.y,
              eos, ev, rem, lem);
  Cons<T> Fl, Fr;
  Fl.d = Ul.Mx;                                                 // :136
  Fr.d = Ur.Mx;
  Fl.Mx = Ul.Mx * Wl.Vx;
  Fr.Mx = Ur.Mx * Wr.Vx;
  Fl.My = Ul.Mx * Wl.Vy;
  Fr.My = Ur.Mx * Wr.Vy;
  Fl.Mz = Ul.Mx * Wl.Vz;
  Fr.Mz = Ur.Mx * Wr.Vz;
  Fl.Mx = Fl.Mx + Wl.P;
  Fr.Mx = Fr.Mx + Wr.P;
  Fl.E = (Ul.E + Wl.P) * Wl.Vx;
  Fr.E = (Ur.E + Wr.P) * Wr.Vx;
  Fl.Mx = Fl.Mx - 0.5 * (Bxi * Bxi - SQR(Wl.By) - SQR(Wl.Bz));
  Fr.Mx = Fr.Mx - 0.5 * (Bxi * Bxi - SQR(Wr.By) - SQR(Wr.Bz));
  Fl.My = Fl.My - Bxi * Wl.By;
  Fr.My = Fr.My - Bxi * Wr.By;
  Fl.Mz = Fl.Mz - Bxi * Wl.Bz;
  Fr.Mz = Fr.Mz - Bxi * Wr.Bz;
  Fl.E = Fl.E + (r.pbl * Wl.Vx - Bxi * (Bxi * Wl.Vx + Wl.By * Wl.Vy + Wl.Bz * Wl.Vz));
  Fr.E = Fr.E + (r.pbr * Wr.Vx - Bxi * (Bxi * Wr.Vx + Wr.By * Wr.Vy + Wr.Bz * Wr.Vz));
  Fl.By = Wl.By * Wl.Vx - Bxi * Wl.Vy;
  Fr.By = Wr.By * Wr.Vx - Bxi * Wr.Vy;
  Fl.Bz = Wl.Bz * Wl.Vx - Bxi * Wl.Vz;
  Fr.Bz = Wr.Bz * Wr.Vx - Bxi * Wr.Vz;
  if (used_hlle) *used_hlle = 0;
  if (pv(ev[0]) >= 0.0) return Fl;                              // :181
  if (pv(ev[NWAVE - 1]) <= 0.0) return Fr;
  const T* pUl = as_arr(Ul);
  const T* pUr = as_arr(Ur);
  T dU[NWAVE], a[NWAVE], u_inter[NWAVE], coeff[NWAVE], p_inter;
  for (int n = 0; n < NWAVE; n++) dU[n] = pUr[n] - pUl[n];      // :200
  for (int n = 0; n < NWAVE; n++) {
    a[n] = T(0.0);
    for (int m = 0; m < NWAVE; m++) a[n] = a[n] + lem[n][m] * dU[m];
  }
  int hlle_flag = 0;                                            // :212
  for (int n = 0; n < NWAVE; n++) u_inter[n] = pUl[n];
  for (int n = 0; n < NWAVE - 1; n++) {
    for (int m = 0; m < NWAVE; m++) u_inter[m] = u_inter[m] + a[n] * rem[m][n];
    if (pv(ev[n + 1]) > pv(ev[n])) {
      if (pv(u_inter[0]) <= 0.0) {
        hlle_flag = 1;
        break;
      }
      p_inter = u_inter[4] - 0.5 * (SQR(u_inter[1]) + SQR(u_inter[2]) + SQR(u_inter[3])) / u_inter[0];
      p_inter = p_inter - 0.5 * (SQR(u_inter[NWAVE - 2]) + SQR(u_inter[NWAVE - 1]) + SQR(Bxi));
      if (pv(p_inter) < 0.0) {
        hlle_flag = 2;
        break;
      }
    }
  }
  if (hlle_flag != 0) {
    if (used_hlle) *used_hlle = hlle_flag;
    return flux_hlle(Ul, Ur, Wl, Wr, Bxi, eos);
  }
  const T* pFl = as_arr(Fl);                                    // :244
  const T* pFr = as_arr(Fr);
  Cons<T> F;
  T* pF = as_arr(F);
  for (int m = 0; m < NWAVE; m++) coeff[m] = 0.5 * MAXT(T(fabs(ev[m])), T(etah)) * a[m];
  for (int n = 0; n < NWAVE; n++) {
    pF[n] = 0.5 * (pFl[n] + pFr[n]);
    for (int m = 0; m < NWAVE; m++) pF[n] = pF[n] - coeff[m] * rem[n][m];
  }
  return F;
}

enum FluxKind { FLUX_ROE = 0, FLUX_HLLE = 1 };

template <class T>
inline Cons<T> riemann(int kind, const Cons<T>& Ul, const Cons<T>& Ur, const Prim<T>& Wl,
                       const Prim<T>& Wr, const T& Bxi, const Eos& eos) {
  return kind == FLUX_HLLE ? flux_hlle(Ul, Ur, Wl, Wr, Bxi, eos)
                           : flux_roe(Ul, Ur, Wl, Wr, Bxi, eos);
}

// ---------------------------------------------------------------------------
// The grid: init_grid.c:25-118 geometry, athena.h:106-120 contents, stored as
// planes [sy][sx] with sx = Nx1 + 2*nghost, sy = Nx2 + 2*nghost.
enum { BC_REFLECT = 1, BC_OUTFLOW = 2, BC_PERIODIC = 4 };

template <class T> struct Grid {
  int Nx1 = 0, Nx2 = 0, is = 0, ie = 0, js = 0, je = 0, sx = 0, sy = 0, idisp = 0, jdisp = 0;
  double x1_0 = 0, x2_0 = 0, dx1 = 0, dx2 = 0, dt = 0, time = 0, cour_no = 0;
  int nstep = 0;
  int bc[4] = {BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC};  // ix1, ox1, ix2, ox2
  int flux = FLUX_ROE;
  Eos eos;
  // conserved variables in Gas order (athena.h:43-52) and interface fields
  std::vector<T> d, M1, M2, M3, E, B1c, B2c, B3c, B1i, B2i, B3i;
  inline int id(int j, int i) const { return j * sx + i; }
  std::vector<T>* plane(int k) {
    std::vector<T>* p[11] = {&d, &M1, &M2, &M3, &E, &B1c, &B2c, &B3c, &B1i, &B2i, &B3i};
    return p[k];
  }
};

template <class T>
void init_grid(Grid<T>& G, int Nx1, int Nx2, double x1min, double x1max, double x2min,
               double x2max, double gamma, double cour_no) {
  G.Nx1 = Nx1;
  G.Nx2 = Nx2;
  G.is = NGHOST;
  G.ie = Nx1 + NGHOST - 1;
  G.js = NGHOST;
  G.je = Nx2 + NGHOST - 1;
  G.dx1 = (x1max - x1min) / (double)Nx1;                        // init_grid.c:60
  G.dx2 = (x2max - x2min) / (double)Nx2;
  G.idisp = -G.is;
  G.jdisp = -G.js;
  G.x1_0 = x1min;
  G.x2_0 = x2min;
  G.sx = Nx1 + 2 * NGHOST;
  G.sy = Nx2 + 2 * NGHOST;
  G.time = 0.0;
  G.nstep = 0;
  G.eos = Eos(gamma);
  G.cour_no = cour_no;
  for (int k = 0; k < 11; k++) G.plane(k)->assign((size_t)G.sx * G.sy, T(0.0));
}

// cc_pos.c:25-31
template <class T>
inline void cc_pos(const Grid<T>& G, int i, int j, double* x1, double* x2) {
  *x1 = G.x1_0 + ((i + G.idisp) + 0.5) * G.dx1;
  *x2 = G.x2_0 + ((j + G.jdisp) + 0.5) * G.dx2;
}

// ---------------------------------------------------------------------------
// Scratch of one integrate_2d call (__integrate_2d__.c:56-70, 573-623)
template <class T> struct Scratch {
  std::vector<Cons<T>> Ul1, Ur1, Ul2, Ur2, F1, F2;
  std::vector<T> B1h, B2h, emf3, emf3cc, dhalf;
  void alloc(size_t n) {
    Ul1.assign(n, Cons<T>()); Ur1.assign(n, Cons<T>());
    Ul2.assign(n, Cons<T>()); Ur2.assign(n, Cons<T>());
    F1.assign(n, Cons<T>()); F2.assign(n, Cons<T>());
    B1h.assign(n, T(0.0)); B2h.assign(n, T(0.0));
    emf3.assign(n, T(0.0)); emf3cc.assign(n, T(0.0)); dhalf.assign(n, T(0.0));
  }
};

// __integrate_2d__.c:632-702
template <class T>
void integrate_emf3_corner(const Grid<T>& G, Scratch<T>& S) {
  const int is = G.is, ie = G.ie, js = G.js, je = G.je;
  const auto& F1 = S.F1;
  const auto& F2 = S.F2;
  const auto& ecc = S.emf3cc;
  auto at = [&](int j, int i) { return G.id(j, i); };
#pragma omp parallel for schedule(static)
  for (int j = js - 1; j <= je + 2; j++) {
    for (int i = is - 1; i <= ie + 2; i++) {
      T emf_l1, emf_r1, emf_l2, emf_r2;
      if (pv(F1[at(j - 1, i)].d) > 0.0) {
        emf_l2 = -F1[at(j - 1, i)].By + (F2[at(j, i - 1)].Bz - ecc[at(j - 1, i - 1)]);
      } else if (pv(F1[at(j - 1, i)].d) < 0.0) {
        emf_l2 = -F1[at(j - 1, i)].By + (F2[at(j, i)].Bz - ecc[at(j - 1, i)]);
      } else {
        emf_l2 = -F1[at(j - 1, i)].By +
                 0.5 * (F2[at(j, i - 1)].Bz - ecc[at(j - 1, i - 1)] + F2[at(j, i)].Bz - ecc[at(j - 1, i)]);
      }
      if (pv(F1[at(j, i)].d) > 0.0) {
        emf_r2 = -F1[at(j, i)].By + (F2[at(j, i - 1)].Bz - ecc[at(j, i - 1)]);
      } else if (pv(F1[at(j, i)].d) < 0.0) {
        emf_r2 = -F1[at(j, i)].By + (F2[at(j, i)].Bz - ecc[at(j, i)]);
      } else {
        emf_r2 = -F1[at(j, i)].By +
                 0.5 * (F2[at(j, i - 1)].Bz - ecc[at(j, i - 1)] + F2[at(j, i)].Bz - ecc[at(j, i)]);
      }
      if (pv(F2[at(j, i - 1)].d) > 0.0) {
        emf_l1 = F2[at(j, i - 1)].Bz + (-F1[at(j - 1, i)].By - ecc[at(j - 1, i - 1)]);
      } else if (pv(F2[at(j, i - 1)].d) < 0.0) {
        emf_l1 = F2[at(j, i - 1)].Bz + (-F1[at(j, i)].By - ecc[at(j, i - 1)]);
      } else {
        emf_l1 = F2[at(j, i - 1)].Bz +
                 0.5 * (-F1[at(j - 1, i)].By - ecc[at(j - 1, i - 1)] - F1[at(j, i)].By - ecc[at(j, i - 1)]);
      }
      if (pv(F2[at(j, i)].d) > 0.0) {
        emf_r1 = F2[at(j, i)].Bz + (-F1[at(j - 1, i)].By - ecc[at(j - 1, i)]);
      } else if (pv(F2[at(j, i)].d) < 0.0) {
        emf_r1 = F2[at(j, i)].Bz + (-F1[at(j, i)].By - ecc[at(j, i)]);
      } else {
        emf_r1 = F2[at(j, i)].Bz +
                 0.5 * (-F1[at(j - 1, i)].By - ecc[at(j - 1, i)] - F1[at(j, i)].By - ecc[at(j, i)]);
      }
      S.emf3[at(j, i)] = 0.25 * (emf_l1 + emf_r1 + emf_l2 + emf_r2);
    }
  }
}

// __integrate_2d__.c:100-537: one CTU + CT step of G by G.dt
template <class T>
void integrate_2d(Grid<T>& G, Scratch<T>& S) {
  const double dtodx1 = G.dt / G.dx1, dtodx2 = G.dt / G.dx2;
  const double hdtodx1 = 0.5 * dtodx1, hdtodx2 = 0.5 * dtodx2;
  const double hdt = 0.5 * G.dt;
  const int is = G.is, ie = G.ie, js = G.js, je = G.je;
  const int il = is - 2, iu = ie + 2, jl = js - 2, ju = je + 2;
  const int sx = G.sx, sy = G.sy, nmax = sx > sy ? sx : sy;
  const Eos& eos = G.eos;
  auto at = [&](int j, int i) { return G.id(j, i); };
  if (S.F1.size() != (size_t)sx * sy) S.alloc((size_t)sx * sy);
  std::vector<T>& d = G.d; std::vector<T>& M1 = G.M1; std::vector<T>& M2 = G.M2;
  std::vector<T>& M3 = G.M3; std::vector<T>& E = G.E; std::vector<T>& B1c = G.B1c;
  std::vector<T>& B2c = G.B2c; std::vector<T>& B3c = G.B3c; std::vector<T>& B1i = G.B1i;
  std::vector<T>& B2i = G.B2i; std::vector<T>& B3i = G.B3i;

  // Step 1a/1b/1e: x1 sweep (:128-180)
#pragma omp parallel
  {
    std::vector<Prim<T>> W(nmax), Wl(nmax), Wr(nmax);
    std::vector<T> Bxc(nmax);
#pragma omp for schedule(static)
    for (int j = jl; j <= ju; j++) {
      for (int i = is - NGHOST; i <= ie + NGHOST; i++) {
        Cons<T> u{d[at(j, i)], M1[at(j, i)], M2[at(j, i)], M3[at(j, i)], E[at(j, i)],
                  B2c[at(j, i)], B3c[at(j, i)]};
        Bxc[i] = B1c[at(j, i)];
        S.B1h[at(j, i)] = B1i[at(j, i)];
        W[i] = cons_to_prim(u, Bxc[i], eos);
      }
      lr_states(W.data(), Bxc.data(), G.dt, dtodx1, is - 1, ie + 1, Wl.data(), Wr.data(), eos);
      for (int i = is - 1; i <= iu; i++) {                      // :156-164
        T src = (M2[at(j, i - 1)] / d[at(j, i - 1)]) * (B1i[at(j, i)] - B1i[at(j, i - 1)]) / G.dx1;
        Wl[i].By = Wl[i].By + hdt * src;
        src = (M2[at(j, i)] / d[at(j, i)]) * (B1i[at(j, i + 1)] - B1i[at(j, i)]) / G.dx1;
        Wr[i].By = Wr[i].By + hdt * src;
      }
      for (int i = is - 1; i <= iu; i++) {                      // :172-180
        Cons<T> ul = prim_to_cons(Wl[i], B1i[at(j, i)], eos);
        Cons<T> ur = prim_to_cons(Wr[i], B1i[at(j, i)], eos);
        S.Ul1[at(j, i)] = ul;
        S.Ur1[at(j, i)] = ur;
        S.F1[at(j, i)] = riemann(G.flux, ul, ur, Wl[i], Wr[i], B1i[at(j, i)], eos);
      }
    }
  }

  // Step 2a/2b/2e: x2 sweep with (d, M2, M3, M1, E, B3c, B1c) (:189-241)
#pragma omp parallel
  {
    std::vector<Prim<T>> W(nmax), Wl(nmax), Wr(nmax);
    std::vector<T> Bxc(nmax);
#pragma omp for schedule(static)
    for (int i = il; i <= iu; i++) {
      for (int j = js - NGHOST; j <= je + NGHOST; j++) {
        Cons<T> u{d[at(j, i)], M2[at(j, i)], M3[at(j, i)], M1[at(j, i)], E[at(j, i)],
                  B3c[at(j, i)], B1c[at(j, i)]};
        Bxc[j] = B2c[at(j, i)];
        S.B2h[at(j, i)] = B2i[at(j, i)];
        W[j] = cons_to_prim(u, Bxc[j], eos);
      }
      lr_states(W.data(), Bxc.data(), G.dt, dtodx2, js - 1, je + 1, Wl.data(), Wr.data(), eos);
      for (int j = js - 1; j <= ju; j++) {                      // :216-224
        T src = (M1[at(j - 1, i)] / d[at(j - 1, i)]) * (B2i[at(j, i)] - B2i[at(j - 1, i)]) / G.dx2;
        Wl[j].Bz = Wl[j].Bz + hdt * src;
        src = (M1[at(j, i)] / d[at(j, i)]) * (B2i[at(j + 1, i)] - B2i[at(j, i)]) / G.dx2;
        Wr[j].Bz = Wr[j].Bz + hdt * src;
      }
      for (int j = js - 1; j <= ju; j++) {                      // :233-241
        Cons<T> ul = prim_to_cons(Wl[j], B2i[at(j, i)], eos);
        Cons<T> ur = prim_to_cons(Wr[j], B2i[at(j, i)], eos);
        S.Ul2[at(j, i)] = ul;
        S.Ur2[at(j, i)] = ur;
        S.F2[at(j, i)] = riemann(G.flux, ul, ur, Wl[j], Wr[j], B2i[at(j, i)], eos);
      }
    }
  }

  // Step 3: cell-centred emf3 at t^n (:249-255)
#pragma omp parallel for schedule(static)
  for (int j = jl; j <= ju; j++)
    for (int i = il; i <= iu; i++)
      S.emf3cc[at(j, i)] = (B1c[at(j, i)] * M2[at(j, i)] - B2c[at(j, i)] * M1[at(j, i)]) / d[at(j, i)];

  // Step 4: corner emf and half-step CT update of the face fields (:263-275)
  integrate_emf3_corner(G, S);
#pragma omp parallel for schedule(static)
  for (int j = js - 1; j <= je + 1; j++) {
    for (int i = is - 1; i <= ie + 1; i++) {
      S.B1h[at(j, i)] = S.B1h[at(j, i)] - hdtodx2 * (S.emf3[at(j + 1, i)] - S.emf3[at(j, i)]);
      S.B2h[at(j, i)] = S.B2h[at(j, i)] + hdtodx1 * (S.emf3[at(j, i + 1)] - S.emf3[at(j, i)]);
    }
    S.B1h[at(j, iu)] = S.B1h[at(j, iu)] - hdtodx2 * (S.emf3[at(j + 1, iu)] - S.emf3[at(j, iu)]);
  }
  for (int i = is - 1; i <= ie + 1; i++)
    S.B2h[at(ju, i)] = S.B2h[at(ju, i)] + hdtodx1 * (S.emf3[at(ju, i + 1)] - S.emf3[at(ju, i)]);

  // Step 5a/5b: transverse x2-flux gradients and MHD sources on x1-face states (:284-334)
#pragma omp parallel for schedule(static)
  for (int j = js - 1; j <= je + 1; j++) {
    for (int i = is - 1; i <= iu; i++) {
      Cons<T>& L = S.Ul1[at(j, i)];
      Cons<T>& R = S.Ur1[at(j, i)];
      const Cons<T>& gl1 = S.F2[at(j + 1, i - 1)];
      const Cons<T>& gl0 = S.F2[at(j, i - 1)];
      const Cons<T>& gr1 = S.F2[at(j + 1, i)];
      const Cons<T>& gr0 = S.F2[at(j, i)];
      L.d = L.d - hdtodx2 * (gl1.d - gl0.d);
      L.Mx = L.Mx - hdtodx2 * (gl1.Mz - gl0.Mz);
      L.My = L.My - hdtodx2 * (gl1.Mx - gl0.Mx);
      L.Mz = L.Mz - hdtodx2 * (gl1.My - gl0.My);
      L.E = L.E - hdtodx2 * (gl1.E - gl0.E);
      L.Bz = L.Bz - hdtodx2 * (gl1.By - gl0.By);
      R.d = R.d - hdtodx2 * (gr1.d - gr0.d);
      R.Mx = R.Mx - hdtodx2 * (gr1.Mz - gr0.Mz);
      R.My = R.My - hdtodx2 * (gr1.Mx - gr0.Mx);
      R.Mz = R.Mz - hdtodx2 * (gr1.My - gr0.My);
      R.E = R.E - hdtodx2 * (gr1.E - gr0.E);
      R.Bz = R.Bz - hdtodx2 * (gr1.By - gr0.By);
      T dbx = B1i[at(j, i)] - B1i[at(j, i - 1)];
      T b1 = B1c[at(j, i - 1)], b2 = B2c[at(j, i - 1)], b3 = B3c[at(j, i - 1)];
      T v3 = M3[at(j, i - 1)] / d[at(j, i - 1)];
      L.Mx = L.Mx + hdtodx1 * b1 * dbx;
      L.My = L.My + hdtodx1 * b2 * dbx;
      L.Mz = L.Mz + hdtodx1 * b3 * dbx;
      L.Bz = L.Bz + hdtodx1 * v3 * dbx;
      L.E = L.E + hdtodx1 * b3 * v3 * dbx;
      dbx = B1i[at(j, i + 1)] - B1i[at(j, i)];
      b1 = B1c[at(j, i)]; b2 = B2c[at(j, i)]; b3 = B3c[at(j, i)];
      v3 = M3[at(j, i)] / d[at(j, i)];
      R.Mx = R.Mx + hdtodx1 * b1 * dbx;
      R.My = R.My + hdtodx1 * b2 * dbx;
      R.Mz = R.Mz + hdtodx1 * b3 * dbx;
      R.Bz = R.Bz + hdtodx1 * v3 * dbx;
      R.E = R.E + hdtodx1 * b3 * v3 * dbx;
    }
  }

  // Step 6a/6b: transverse x1-flux gradients and MHD sources on x2-face states (:343-393)
#pragma omp parallel for schedule(static)
  for (int j = js - 1; j <= ju; j++) {
    for (int i = is - 1; i <= ie + 1; i++) {
      Cons<T>& L = S.Ul2[at(j, i)];
      Cons<T>& R = S.Ur2[at(j, i)];
      const Cons<T>& fl1 = S.F1[at(j - 1, i + 1)];
      const Cons<T>& fl0 = S.F1[at(j - 1, i)];
      const Cons<T>& fr1 = S.F1[at(j, i + 1)];
      const Cons<T>& fr0 = S.F1[at(j, i)];
      L.d = L.d - hdtodx1 * (fl1.d - fl0.d);
      L.Mx = L.Mx - hdtodx1 * (fl1.My - fl0.My);
      L.My = L.My - hdtodx1 * (fl1.Mz - fl0.Mz);
      L.Mz = L.Mz - hdtodx1 * (fl1.Mx - fl0.Mx);
      L.E = L.E - hdtodx1 * (fl1.E - fl0.E);
      L.By = L.By - hdtodx1 * (fl1.Bz - fl0.Bz);
      R.d = R.d - hdtodx1 * (fr1.d - fr0.d);
      R.Mx = R.Mx - hdtodx1 * (fr1.My - fr0.My);
      R.My = R.My - hdtodx1 * (fr1.Mz - fr0.Mz);
      R.Mz = R.Mz - hdtodx1 * (fr1.Mx - fr0.Mx);
      R.E = R.E - hdtodx1 * (fr1.E - fr0.E);
      R.By = R.By - hdtodx1 * (fr1.Bz - fr0.Bz);
      T dby = B2i[at(j, i)] - B2i[at(j - 1, i)];
      T b1 = B1c[at(j - 1, i)], b2 = B2c[at(j - 1, i)], b3 = B3c[at(j - 1, i)];
      T v3 = M3[at(j - 1, i)] / d[at(j - 1, i)];
      L.Mz = L.Mz + hdtodx2 * b1 * dby;
      L.Mx = L.Mx + hdtodx2 * b2 * dby;
      L.My = L.My + hdtodx2 * b3 * dby;
      L.By = L.By + hdtodx2 * v3 * dby;
      L.E = L.E + hdtodx2 * b3 * v3 * dby;
      dby = B2i[at(j + 1, i)] - B2i[at(j, i)];
      b1 = B1c[at(j, i)]; b2 = B2c[at(j, i)]; b3 = B3c[at(j, i)];
      v3 = M3[at(j, i)] / d[at(j, i)];
      R.Mz = R.Mz + hdtodx2 * b1 * dby;
      R.Mx = R.Mx + hdtodx2 * b2 * dby;
      R.My = R.My + hdtodx2 * b3 * dby;
      R.By = R.By + hdtodx2 * v3 * dby;
      R.E = R.E + hdtodx2 * b3 * v3 * dby;
    }
  }

  // Step 7: d and emf3 at t^{n+1/2} (:402-431)
#pragma omp parallel for schedule(static)
  for (int j = js - 1; j <= je + 1; j++) {
    for (int i = is - 1; i <= ie + 1; i++) {
      S.dhalf[at(j, i)] = d[at(j, i)] - hdtodx1 * (S.F1[at(j, i + 1)].d - S.F1[at(j, i)].d) -
                          hdtodx2 * (S.F2[at(j + 1, i)].d - S.F2[at(j, i)].d);
      T m1 = M1[at(j, i)] - hdtodx1 * (S.F1[at(j, i + 1)].Mx - S.F1[at(j, i)].Mx) -
             hdtodx2 * (S.F2[at(j + 1, i)].Mz - S.F2[at(j, i)].Mz);
      T m2 = M2[at(j, i)] - hdtodx1 * (S.F1[at(j, i + 1)].My - S.F1[at(j, i)].My) -
             hdtodx2 * (S.F2[at(j + 1, i)].Mx - S.F2[at(j, i)].Mx);
      T b1 = 0.5 * (S.B1h[at(j, i)] + S.B1h[at(j, i + 1)]);
      T b2 = 0.5 * (S.B2h[at(j, i)] + S.B2h[at(j + 1, i)]);
      S.emf3cc[at(j, i)] = (b1 * m2 - b2 * m1) / S.dhalf[at(j, i)];
    }
  }

  // Step 8b: corrector x1 fluxes (:438-447)
#pragma omp parallel for schedule(static)
  for (int j = js - 1; j <= je + 1; j++) {
    for (int i = is; i <= ie + 1; i++) {
      const T bx = S.B1h[at(j, i)];
      Prim<T> wl = cons_to_prim(S.Ul1[at(j, i)], bx, eos);
      Prim<T> wr = cons_to_prim(S.Ur1[at(j, i)], bx, eos);
      S.F1[at(j, i)] = riemann(G.flux, S.Ul1[at(j, i)], S.Ur1[at(j, i)], wl, wr, bx, eos);
    }
  }
  // Step 8c: corrector x2 fluxes (:455-464)
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je + 1; j++) {
    for (int i = is - 1; i <= ie + 1; i++) {
      const T bx = S.B2h[at(j, i)];
      Prim<T> wl = cons_to_prim(S.Ul2[at(j, i)], bx, eos);
      Prim<T> wr = cons_to_prim(S.Ur2[at(j, i)], bx, eos);
      S.F2[at(j, i)] = riemann(G.flux, S.Ul2[at(j, i)], S.Ur2[at(j, i)], wl, wr, bx, eos);
    }
  }

  // Step 9: corner emf at t^{n+1/2} and full-step CT update (:473-485)
  integrate_emf3_corner(G, S);
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; j++) {
    for (int i = is; i <= ie; i++) {
      B1i[at(j, i)] = B1i[at(j, i)] - dtodx2 * (S.emf3[at(j + 1, i)] - S.emf3[at(j, i)]);
      B2i[at(j, i)] = B2i[at(j, i)] + dtodx1 * (S.emf3[at(j, i + 1)] - S.emf3[at(j, i)]);
    }
    B1i[at(j, ie + 1)] = B1i[at(j, ie + 1)] - dtodx2 * (S.emf3[at(j + 1, ie + 1)] - S.emf3[at(j, ie + 1)]);
  }
  for (int i = is; i <= ie; i++)
    B2i[at(je + 1, i)] = B2i[at(je + 1, i)] + dtodx1 * (S.emf3[at(je + 1, i + 1)] - S.emf3[at(je + 1, i)]);

  // Step 11a/11b: conservative update, x1 then x2 (:491-518); Step 13 (:526-534)
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; j++) {
    for (int i = is; i <= ie; i++) {
      const Cons<T>& f0 = S.F1[at(j, i)];
      const Cons<T>& f1 = S.F1[at(j, i + 1)];
      d[at(j, i)] = d[at(j, i)] - dtodx1 * (f1.d - f0.d);
      M1[at(j, i)] = M1[at(j, i)] - dtodx1 * (f1.Mx - f0.Mx);
      M2[at(j, i)] = M2[at(j, i)] - dtodx1 * (f1.My - f0.My);
      M3[at(j, i)] = M3[at(j, i)] - dtodx1 * (f1.Mz - f0.Mz);
      E[at(j, i)] = E[at(j, i)] - dtodx1 * (f1.E - f0.E);
      B2c[at(j, i)] = B2c[at(j, i)] - dtodx1 * (f1.By - f0.By);
      B3c[at(j, i)] = B3c[at(j, i)] - dtodx1 * (f1.Bz - f0.Bz);
      const Cons<T>& g0 = S.F2[at(j, i)];
      const Cons<T>& g1 = S.F2[at(j + 1, i)];
      d[at(j, i)] = d[at(j, i)] - dtodx2 * (g1.d - g0.d);
      M1[at(j, i)] = M1[at(j, i)] - dtodx2 * (g1.Mz - g0.Mz);
      M2[at(j, i)] = M2[at(j, i)] - dtodx2 * (g1.Mx - g0.Mx);
      M3[at(j, i)] = M3[at(j, i)] - dtodx2 * (g1.My - g0.My);
      E[at(j, i)] = E[at(j, i)] - dtodx2 * (g1.E - g0.E);
      B3c[at(j, i)] = B3c[at(j, i)] - dtodx2 * (g1.By - g0.By);
      B1c[at(j, i)] = B1c[at(j, i)] - dtodx2 * (g1.Bz - g0.Bz);
    }
  }
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; j++) {
    for (int i = is; i <= ie; i++) {
      B1c[at(j, i)] = 0.5 * (B1i[at(j, i)] + B1i[at(j, i + 1)]);
      B2c[at(j, i)] = 0.5 * (B2i[at(j, i)] + B2i[at(j + 1, i)]);
      B3i[at(j, i)] = B3c[at(j, i)];
    }
  }
}

inline double pressure_floor(double p) { return (p > TINY_NUMBER) ? p : TINY_NUMBER; }

// new_dt.c:25-103 (the fast-speed estimate is computed and then replaced by the
// sound speed from p = Gamma_1*(E - d v^2/2), :68-77; only that path is kept)
template <class T>
double max_inverse_dt(const Grid<T>& G) {
  using std::fabs;
  using std::sqrt;
  double max_dti = 0.0;
#pragma omp parallel for reduction(max : max_dti) schedule(static)
  for (int j = G.js; j <= G.je; j++) {
    for (int i = G.is; i <= G.ie; i++) {
      const int k = G.id(j, i);
      double di = 1.0 / pv(G.d[k]);
      double v1 = pv(G.M1[k]) * di;
      double v2 = pv(G.M2[k]) * di;
      double v3 = pv(G.M3[k]) * di;
      double qsq = v1 * v1 + v2 * v2 + v3 * v3;
      double p = pressure_floor(G.eos.g1 * (pv(G.E[k]) - 0.5 * pv(G.d[k]) * qsq));
      double asq = G.eos.gamma * p * di;
      double cf1sq = asq, cf2sq = asq;
      if (G.Nx1 > 1) {
        double c = (fabs(v1) + sqrt(cf1sq)) / G.dx1;
        max_dti = (max_dti > c) ? max_dti : c;
      }
      if (G.Nx2 > 1) {
        double c = (fabs(v2) + sqrt(cf2sq)) / G.dx2;
        max_dti = (max_dti > c) ? max_dti : c;
      }
    }
  }
  return max_dti;
}

template <class T>
void new_dt(Grid<T>& G) {
  double max_dti = max_inverse_dt(G);
  if (G.nstep == 0) {                                           // :93-97
    G.dt = G.cour_no / max_dti;
  } else {
    double a = 2.0 * G.dt, b = G.cour_no / max_dti;
    G.dt = (a < b) ? a : b;
  }
}

// ---------------------------------------------------------------------------
// set_bvals.c:114-145: x1 boundaries, then x2 boundaries (corners come out
// right); reflect :312-545, outflow :553-786, periodic :793-1031.
template <class T>
void set_bvals(Grid<T>& G) {
  const int is = G.is, ie = G.ie, js = G.js, je = G.je;
  auto at = [&](int j, int i) { return G.id(j, i); };
  std::vector<T>* U[8] = {&G.d, &G.M1, &G.M2, &G.M3, &G.E, &G.B1c, &G.B2c, &G.B3c};
  auto copy_cell = [&](int jd, int id, int js_, int is_) {
    for (int k = 0; k < 8; k++) (*U[k])[at(jd, id)] = (*U[k])[at(js_, is_)];
  };
  // x1 inner (:793-848 periodic, :553-605 outflow, :312-372 reflect)
  if (G.Nx1 > 1) {
    const int bi = G.bc[0], bo = G.bc[1];
    const int ju = (G.Nx2 > 1) ? je + 1 : je;
    for (int j = js; j <= je; j++)
      for (int i = 1; i <= NGHOST; i++) {
        if (bi == BC_PERIODIC) copy_cell(j, is - i, j, ie - (i - 1));
        else if (bi == BC_OUTFLOW) copy_cell(j, is - i, j, is);
        else { copy_cell(j, is - i, j, is + (i - 1)); G.M1[at(j, is - i)] = -G.M1[at(j, is - i)]; }
      }
    for (int j = js; j <= je; j++)
      for (int i = 1; i <= NGHOST; i++) {
        if (bi == BC_PERIODIC) G.B1i[at(j, is - i)] = G.B1i[at(j, ie - (i - 1))];
        else if (bi == BC_OUTFLOW) G.B1i[at(j, is - i)] = G.B1i[at(j, is)];
        else G.B1i[at(j, is - i)] = G.B1i[at(j, is + (i - 1))];
      }
    for (int j = js; j <= ju; j++)
      for (int i = 1; i <= NGHOST; i++) {
        if (bi == BC_PERIODIC) G.B2i[at(j, is - i)] = G.B2i[at(j, ie - (i - 1))];
        else if (bi == BC_OUTFLOW) G.B2i[at(j, is - i)] = G.B2i[at(j, is)];
        else G.B2i[at(j, is - i)] = G.B2i[at(j, is + (i - 1))];
      }
    for (int j = js; j <= je; j++)
      for (int i = 1; i <= NGHOST; i++) {
        if (bi == BC_PERIODIC) G.B3i[at(j, is - i)] = G.B3i[at(j, ie - (i - 1))];
        else if (bi == BC_OUTFLOW) G.B3i[at(j, is - i)] = G.B3i[at(j, is)];
        else G.B3i[at(j, is - i)] = G.B3i[at(j, is + (i - 1))];
      }
    // x1 outer; i = ie+1 is not a boundary value for B1i
    for (int j = js; j <= je; j++)
      for (int i = 1; i <= NGHOST; i++) {
        if (bo == BC_PERIODIC) copy_cell(j, ie + i, j, is + (i - 1));
        else if (bo == BC_OUTFLOW) copy_cell(j, ie + i, j, ie);
        else { copy_cell(j, ie + i, j, ie - (i - 1)); G.M1[at(j, ie + i)] = -G.M1[at(j, ie + i)]; }
      }
    for (int j = js; j <= je; j++)
      for (int i = 2; i <= NGHOST; i++) {
        if (bo == BC_PERIODIC) G.B1i[at(j, ie + i)] = G.B1i[at(j, is + (i - 1))];
        else if (bo == BC_OUTFLOW) G.B1i[at(j, ie + i)] = G.B1i[at(j, ie)];
        else G.B1i[at(j, ie + i)] = G.B1i[at(j, ie - (i - 2))];
      }
    for (int j = js; j <= ju; j++)
      for (int i = 1; i <= NGHOST; i++) {
        if (bo == BC_PERIODIC) G.B2i[at(j, ie + i)] = G.B2i[at(j, is + (i - 1))];
        else if (bo == BC_OUTFLOW) G.B2i[at(j, ie + i)] = G.B2i[at(j, ie)];
        else G.B2i[at(j, ie + i)] = G.B2i[at(j, ie - (i - 1))];
      }
    for (int j = js; j <= je; j++)
      for (int i = 1; i <= NGHOST; i++) {
        if (bo == BC_PERIODIC) G.B3i[at(j, ie + i)] = G.B3i[at(j, is + (i - 1))];
        else if (bo == BC_OUTFLOW) G.B3i[at(j, ie + i)] = G.B3i[at(j, ie)];
        else G.B3i[at(j, ie + i)] = G.B3i[at(j, ie - (i - 1))];
      }
  }
  if (G.Nx2 > 1) {
    const int bi = G.bc[2], bo = G.bc[3];
    const int il = (G.Nx1 > 1) ? is - NGHOST : is, iu = (G.Nx1 > 1) ? ie + NGHOST : ie;
    // x2 inner
    for (int j = 1; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bi == BC_PERIODIC) copy_cell(js - j, i, je - (j - 1), i);
        else if (bi == BC_OUTFLOW) copy_cell(js - j, i, js, i);
        else { copy_cell(js - j, i, js + (j - 1), i); G.M2[at(js - j, i)] = -G.M2[at(js - j, i)]; }
      }
    for (int j = 1; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bi == BC_PERIODIC) G.B1i[at(js - j, i)] = G.B1i[at(je - (j - 1), i)];
        else if (bi == BC_OUTFLOW) G.B1i[at(js - j, i)] = G.B1i[at(js, i)];
        else G.B1i[at(js - j, i)] = G.B1i[at(js + (j - 1), i)];
      }
    for (int j = 1; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bi == BC_PERIODIC) G.B2i[at(js - j, i)] = G.B2i[at(je - (j - 1), i)];
        else if (bi == BC_OUTFLOW) G.B2i[at(js - j, i)] = G.B2i[at(js, i)];
        else G.B2i[at(js - j, i)] = G.B2i[at(js + (j - 1), i)];
      }
    for (int j = 1; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bi == BC_PERIODIC) G.B3i[at(js - j, i)] = G.B3i[at(je - (j - 1), i)];
        else if (bi == BC_OUTFLOW) G.B3i[at(js - j, i)] = G.B3i[at(js, i)];
        else G.B3i[at(js - j, i)] = G.B3i[at(js + (j - 1), i)];
      }
    // x2 outer; j = je+1 is not a boundary value for B2i
    for (int j = 1; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bo == BC_PERIODIC) copy_cell(je + j, i, js + (j - 1), i);
        else if (bo == BC_OUTFLOW) copy_cell(je + j, i, je, i);
        else { copy_cell(je + j, i, je - (j - 1), i); G.M2[at(je + j, i)] = -G.M2[at(je + j, i)]; }
      }
    for (int j = 1; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bo == BC_PERIODIC) G.B1i[at(je + j, i)] = G.B1i[at(js + (j - 1), i)];
        else if (bo == BC_OUTFLOW) G.B1i[at(je + j, i)] = G.B1i[at(je, i)];
        else G.B1i[at(je + j, i)] = G.B1i[at(je - (j - 1), i)];
      }
    for (int j = 2; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bo == BC_PERIODIC) G.B2i[at(je + j, i)] = G.B2i[at(js + (j - 1), i)];
        else if (bo == BC_OUTFLOW) G.B2i[at(je + j, i)] = G.B2i[at(je, i)];
        else G.B2i[at(je + j, i)] = G.B2i[at(je - (j - 2), i)];
      }
    for (int j = 1; j <= NGHOST; j++)
      for (int i = il; i <= iu; i++) {
        if (bo == BC_PERIODIC) G.B3i[at(je + j, i)] = G.B3i[at(js + (j - 1), i)];
        else if (bo == BC_OUTFLOW) G.B3i[at(je + j, i)] = G.B3i[at(je, i)];
        else G.B3i[at(je + j, i)] = G.B3i[at(je - (j - 1), i)];
      }
  }
}

// ---------------------------------------------------------------------------
// Problem generators (host libm, as the reference). blast.c:25-72
template <class T>
void pgen_blast(Grid<T>& G, double rin, double pa, double prat, double b0, double angle_deg,
                double cx = 0.0, double cy = 0.0) {
  using std::cos;
  using std::sin;
  using std::sqrt;
  const double theta = (PI / 180.0) * angle_deg;
  const double da = 1.0, ua = 0.0, va = 0.0, wa = 0.0;
  const double bxa = b0 * cos(theta), bya = b0 * sin(theta), bza = 0.0;
  for (int j = G.js; j <= G.je; j++)
    for (int i = G.is; i <= G.ie; i++) {
      const int k = G.id(j, i);
      G.d[k] = T(da);
      G.M1[k] = T(da * ua);
      G.M2[k] = T(da * va);
      G.M3[k] = T(da * wa);
      G.B1i[k] = T(bxa);
      G.B2i[k] = T(bya);
      G.B3i[k] = T(bza);
      G.B1c[k] = T(bxa);
      G.B2c[k] = T(bya);
      G.B3c[k] = T(bza);
      if (i == G.ie && G.ie > G.is) G.B1i[G.id(j, i + 1)] = T(bxa);
      if (j == G.je && G.je > G.js) G.B2i[G.id(j + 1, i)] = T(bya);
      double x1, x2;
      cc_pos(G, i, j, &x1, &x2);
      // blast.c:64 centres the blast at the origin; (cx, cy) offsets it (cx = cy = 0
      // reproduces the reference: x - 0 == x)
      const double rx = x1 - cx, ry = x2 - cy;
      double rad = sqrt(rx * rx + ry * ry);
      double pressure = pa;
      if (rad < rin) pressure = prat * pa;
      G.E[k] = T(pressure / G.eos.g1 + 0.5 * (bxa * bxa + bya * bya + bza * bza) +
                 0.5 * da * (ua * ua + va * va + wa * wa));
    }
}

// field_loop.c:40-137 (iprob = 1: loop in the x1-x2 plane)
template <class T>
void pgen_field_loop(Grid<T>& G, double rad, double amp, double vflow, int iprob) {
  using std::sqrt;
  const int is = G.is, ie = G.ie, js = G.js, je = G.je;
  std::vector<double> az((size_t)G.sx * G.sy, 0.0);
  for (int j = js; j <= je + 1; j++)
    for (int i = is; i <= ie + 1; i++) {
      double x1c, x2c;
      cc_pos(G, i, j, &x1c, &x2c);
      double x1f = x1c - 0.5 * G.dx1;
      double x2f = x2c - 0.5 * G.dx2;
      if (iprob == 1) {
        if ((x1f * x1f + x2f * x2f) < rad * rad) az[G.id(j, i)] = amp * (rad - sqrt(x1f * x1f + x2f * x2f));
      }
    }
  const double x1size = G.dx1 * (double)G.Nx1;
  const double x2size = G.dx2 * (double)G.Nx2;
  const double diag = sqrt(x1size * x1size + x2size * x2size);
  for (int j = js; j <= je; j++)
    for (int i = is; i <= ie; i++) {
      const int k = G.id(j, i);
      G.d[k] = T(1.0);
      G.M1[k] = T(1.0 * vflow * x1size / diag);
      G.M2[k] = T(1.0 * vflow * x2size / diag);
      G.M3[k] = T(0.0);
      G.B1i[k] = T((az[G.id(j + 1, i)] - az[k]) / G.dx2);
      G.B2i[k] = T(-(az[G.id(j, i + 1)] - az[k]) / G.dx1);
      G.B3i[k] = T(0.0);
    }
  for (int j = js; j <= je; j++)
    G.B1i[G.id(j, ie + 1)] = T((az[G.id(j + 1, ie + 1)] - az[G.id(j, ie + 1)]) / G.dx2);
  for (int i = is; i <= ie; i++)
    G.B2i[G.id(je + 1, i)] = T(-(az[G.id(je + 1, i + 1)] - az[G.id(je + 1, i)]) / G.dx1);
  for (int j = js; j <= je; j++)
    for (int i = is; i <= ie; i++) {
      const int k = G.id(j, i);
      G.B1c[k] = 0.5 * (G.B1i[k] + G.B1i[G.id(j, i + 1)]);
      G.B2c[k] = 0.5 * (G.B2i[k] + G.B2i[G.id(j + 1, i)]);
      G.B3c[k] = G.B3i[k];
      G.E[k] = 1.0 / G.eos.g1 + 0.5 * (SQR(G.B1c[k]) + SQR(G.B2c[k]) + SQR(G.B3c[k])) +
               0.5 * (SQR(G.M1[k]) + SQR(G.M2[k]) + SQR(G.M3[k])) / G.d[k];
    }
}

// cpaw2d.c:56-185 and the vector potential A3 (cpaw2d.c:311-315). Returns the
// wavelength; Soln (the initial state, Gas order) is kept for the L1 error.
struct CpawParams { double sin_a, cos_a, b_par, b_perp, k_par, lambda; };

inline double cpaw_A3(const CpawParams& c, double x1, double x2) {
  return c.b_par * (x1 * c.sin_a - x2 * c.cos_a) -
         (c.b_perp / c.k_par) * std::cos(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
}

template <class T>
CpawParams pgen_cpaw2d(Grid<T>& G, double angle, double b_par, double b_perp, double pres,
                       double v_par, std::vector<double>* soln /* 8 planes */) {
  using std::atan;
  using std::cos;
  using std::sin;
  using std::sqrt;
  CpawParams c;
  const double dx1 = G.dx1, dx2 = G.dx2;
  double lambda;
  if (angle == 0.0) {
    c.sin_a = 0.0; c.cos_a = 1.0; lambda = G.Nx1 * G.dx1;
  } else if (angle == 90.0) {
    c.sin_a = 1.0; c.cos_a = 0.0; lambda = G.Nx2 * G.dx2;
  } else {
    if ((G.Nx1 * G.dx1) == (G.Nx2 * G.dx2)) {
      c.cos_a = c.sin_a = sqrt(0.5);
    } else {
      angle = atan((double)(G.Nx1 * G.dx1) / (G.Nx2 * G.dx2));
      c.sin_a = sin(angle);
      c.cos_a = cos(angle);
    }
    if (c.cos_a >= c.sin_a) lambda = G.Nx1 * G.dx1 * c.cos_a;
    else lambda = G.Nx2 * G.dx2 * c.sin_a;
  }
  c.k_par = 2.0 * PI / lambda;
  c.b_par = b_par;
  c.b_perp = b_perp;
  c.lambda = lambda;
  const double den = 1.0;
  const double v_perp = b_perp / sqrt((double)den);
  for (int j = G.js; j <= G.je + 1; j++)
    for (int i = G.is; i <= G.ie + 1; i++) {
      double x1, x2;
      cc_pos(G, i, j, &x1, &x2);
      double cs = cos(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
      x1 = x1 - 0.5 * G.dx1;
      x2 = x2 - 0.5 * G.dx2;
      const int k = G.id(j, i);
      G.B1i[k] = T(-(cpaw_A3(c, x1, x2 + dx2) - cpaw_A3(c, x1, x2)) / dx2);
      G.B2i[k] = T((cpaw_A3(c, x1 + dx1, x2) - cpaw_A3(c, x1, x2)) / dx1);
      G.B3i[k] = T(b_perp * cs);
    }
  if (soln) {
    soln->assign((size_t)8 * G.sx * G.sy, 0.0);
  }
  const size_t np = (size_t)G.sx * G.sy;
  for (int j = G.js; j <= G.je; j++)
    for (int i = G.is; i <= G.ie; i++) {
      double x1, x2;
      cc_pos(G, i, j, &x1, &x2);
      double sn = sin(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
      double cs = cos(c.k_par * (x1 * c.cos_a + x2 * c.sin_a));
      const int k = G.id(j, i);
      double sd = den;
      double s1 = den * (v_par * c.cos_a - v_perp * sn * c.sin_a);
      double s2 = den * (v_par * c.sin_a + v_perp * sn * c.cos_a);
      double s3 = den * v_perp * cs;
      double b1 = 0.5 * (pv(G.B1i[k]) + pv(G.B1i[G.id(j, i + 1)]));
      double b2 = 0.5 * (pv(G.B2i[k]) + pv(G.B2i[G.id(j + 1, i)]));
      double b3 = b_perp * cs;
      double e = pres / G.eos.g1 + 0.5 * (SQR(b1) + SQR(b2) + SQR(b3)) +
                 0.5 * (SQR(s1) + SQR(s2) + SQR(s3)) / den;
      G.d[k] = T(sd); G.M1[k] = T(s1); G.M2[k] = T(s2); G.M3[k] = T(s3);
      G.B1c[k] = T(b1); G.B2c[k] = T(b2); G.B3c[k] = T(b3); G.E[k] = T(e);
      if (soln) {
        double v[8] = {sd, s1, s2, s3, e, b1, b2, b3};
        for (int q = 0; q < 8; q++) (*soln)[q * np + k] = v[q];
      }
    }
  return c;
}

// cpaw2d.c:222-302: L1 error per variable (Gas order d,M1,M2,M3,E,B1c,B2c,B3c)
// divided by Nx1*Nx2, and the RMS of the summed errors.
template <class T>
void cpaw2d_l1(const Grid<T>& G, const std::vector<double>& soln, double out[9]) {
  using std::fabs;
  using std::sqrt;
  const size_t np = (size_t)G.sx * G.sy;
  double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const std::vector<T>* U[8] = {&G.d, &G.M1, &G.M2, &G.M3, &G.E, &G.B1c, &G.B2c, &G.B3c};
  for (int j = G.js; j <= G.je; j++)
    for (int i = G.is; i <= G.ie; i++) {
      const int k = G.id(j, i);
      for (int q = 0; q < 8; q++) e[q] += fabs(pv((*U[q])[k]) - soln[q * np + k]);
    }
  // accumulation order of :255-258: d, M1, M2, M3, B1c, B2c, B3c, E
  double rms = SQR(e[0]) + SQR(e[1]) + SQR(e[2]) + SQR(e[3]);
  rms += SQR(e[5]) + SQR(e[6]) + SQR(e[7]);
  rms += SQR(e[4]);
  const int n = (G.ie - G.is + 1) * (G.je - G.js + 1);
  out[0] = sqrt(rms) / (double)n;
  for (int q = 0; q < 8; q++) out[1 + q] = e[q] / (double)n;
}

}  // namespace ora
