// adjoint_host.cpp -- host evaluation of the hand-written adjoint routines of
// adjoint_device.hpp (the same source the adjoint kernels use), exported for
// checking them one routine at a time against the tangent-linear model.
#include <cstring>

#include "adjoint_device.hpp"
#include "athmi_adjoint.h"

using namespace athmi;
using namespace athmi::adj;

namespace {
EosC eos_of(double gamma) { return EosC{gamma, gamma - 1.0, gamma - 2.0}; }
Cons7 ld_c(const double* a) { return Cons7{a[0], a[1], a[2], a[3], a[4], a[5], a[6]}; }
Prim7 ld_p(const double* a) { return Prim7{a[0], a[1], a[2], a[3], a[4], a[5], a[6]}; }
void st(const Cons7& c, double* a) { std::memcpy(a, &c, sizeof(c)); }
void st(const Prim7& c, double* a) { std::memcpy(a, &c, sizeof(c)); }
}  // namespace

extern "C" {

void athmi_vjp_cons_to_prim(double gamma, const double* u, double bx, const double* wb,
                            double* ub, double* bxb) {
  Cons7 U = zero_cons();
  double b = 0.0;
  cons_to_prim_b(ld_c(u), bx, eos_of(gamma), ld_p(wb), U, b);
  st(U, ub);
  *bxb = b;
}

void athmi_vjp_prim_to_cons(double gamma, const double* w, double bx, const double* ub,
                            double* wb, double* bxb) {
  Prim7 W = zero_prim();
  double b = 0.0;
  prim_to_cons_b(ld_p(w), bx, eos_of(gamma), ld_c(ub), W, b);
  st(W, wb);
  *bxb = b;
}

void athmi_vjp_esys_prim(double gamma, const double* in, const double* Rb, const double* Lb,
                         const double* evb, double* inb) {
  EsysPrimF E;
  E.run(in[0], in[1], in[2], in[3], in[4], in[5], gamma);
  for (int k = 0; k < 6; k++) inb[k] = 0.0;
  E.back([&](int i, int j) { return Rb[7 * i + j]; }, [&](int i, int j) { return Lb[7 * i + j]; },
         evb, inb);
}

void athmi_vjp_esys_roe(double gamma, const double* in, const double* Rb, const double* Lb,
                        const double* evb, double* inb, int vectors) {
  const EosC eos = eos_of(gamma);
  for (int k = 0; k < 10; k++) inb[k] = 0.0;
  if (vectors) {
    EsysRoeF<true> E;
    E.run(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9], eos);
    E.back([&](int i, int j) { return Rb[7 * i + j]; },
           [&](int i, int j) { return Lb[7 * i + j]; }, evb, inb);
  } else {
    EsysRoeF<false> E;
    E.run(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9], eos);
    E.back(ZeroMat(), ZeroMat(), evb, inb);
  }
}

int athmi_vjp_flux(double gamma, int kind, const double* ul, const double* ur, const double* wl,
                   const double* wr, double bx, const double* fb, double* ulb, double* urb,
                   double* wlb, double* wrb, double* bxb) {
  Cons7 Ulb = zero_cons(), Urb = zero_cons();
  Prim7 Wlb = zero_prim(), Wrb = zero_prim();
  double b = 0.0;
  int flag = 0;
  if (kind == kFluxHlle)
    flux_hlle_b(ld_c(ul), ld_c(ur), ld_p(wl), ld_p(wr), bx, eos_of(gamma), ld_c(fb), Ulb, Urb,
                Wlb, Wrb, b);
  else
    flag = flux_roe_b(ld_c(ul), ld_c(ur), ld_p(wl), ld_p(wr), bx, eos_of(gamma), ld_c(fb), Ulb,
                      Urb, Wlb, Wrb, b);
  st(Ulb, ulb);
  st(Urb, urb);
  st(Wlb, wlb);
  st(Wrb, wrb);
  *bxb = b;
  return flag;
}

void athmi_vjp_plm(double gamma, const double* wm, const double* w0, const double* wp, double bxc,
                   double dtodx, const double* out_lb, const double* out_rb, double* out_l,
                   double* out_r, double* wmb, double* w0b, double* wpb, double* bxcb) {
  PlmF P;
  P.run(ld_p(wm), ld_p(w0), ld_p(wp), bxc, dtodx, eos_of(gamma));
  for (int k = 0; k < 7; k++) {
    out_l[k] = P.out_l[k];
    out_r[k] = P.out_r[k];
    wmb[k] = w0b[k] = wpb[k] = 0.0;
  }
  double b = 0.0;
  P.back(out_lb, out_rb, wmb, w0b, wpb, b);
  *bxcb = b;
}

}  // extern "C"
