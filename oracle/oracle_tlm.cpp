// oracle_tlm.cpp -- TEST INFRASTRUCTURE ONLY: the oracle instantiated on dual
// numbers, i.e. the forward-mode tangent-linear model (TLM) of the reference
// routines it restates (esystem_prim.c, esystem_roe.c, flux_roe.c,
// flux_hlle.c, convert_var.c, lr_states_plm.c, __integrate_2d__.c,
// set_bvals.c). The adjoint kernels are checked against it by dot products:
// <ybar, J xdot> from here equals <J^T ybar, xdot> from the adjoint.
#include <omp.h>

#include <cstring>
#include <vector>

#include "dual.hpp"
#include "counted.hpp"
#include "athena_oracle.hpp"

using namespace ora;

namespace ora {
BranchLog& branch_log() {
  static BranchLog L;
  return L;
}
}  // namespace ora

namespace {
template <int N>
void pack(Dual* out, const double* v, const double* t) {
  for (int k = 0; k < N; k++) out[k] = Dual(v[k], t[k]);
}
template <int N>
void unpack(const Dual* in, double* v, double* t) {
  for (int k = 0; k < N; k++) {
    v[k] = in[k].v;
    t[k] = in[k].t;
  }
}
}  // namespace

extern "C" {

void ora_tlm_cons_to_prim(double gamma, const double* u, const double* du, double bx, double dbx,
                          double* w, double* dw) {
  Cons<Dual> U;
  pack<7>(as_arr(U), u, du);
  Prim<Dual> W = cons_to_prim(U, Dual(bx, dbx), Eos(gamma));
  unpack<7>(as_arr(W), w, dw);
}

void ora_tlm_prim_to_cons(double gamma, const double* w, const double* dw, double bx, double dbx,
                          double* u, double* du) {
  Prim<Dual> W;
  pack<7>(as_arr(W), w, dw);
  Cons<Dual> U = prim_to_cons(W, Dual(bx, dbx), Eos(gamma));
  unpack<7>(as_arr(U), u, du);
}

void ora_tlm_esys_prim(double gamma, const double* in, const double* din, double* ev, double* dev,
                       double* R, double* dR, double* L, double* dL) {
  Dual x[6], e[7], r[7][7], l[7][7];
  pack<6>(x, in, din);
  esys_prim<Dual>(x[0], x[1], x[2], x[3], x[4], x[5], Eos(gamma), e, r, l);
  unpack<7>(e, ev, dev);
  unpack<49>(&r[0][0], R, dR);
  unpack<49>(&l[0][0], L, dL);
}

void ora_tlm_esys_roe(double gamma, const double* in, const double* din, double* ev, double* dev,
                      double* R, double* dR, double* L, double* dL) {
  Dual x[10], e[7], r[7][7], l[7][7];
  pack<10>(x, in, din);
  esys_roe<Dual>(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], Eos(gamma), e, r,
                 l);
  unpack<7>(e, ev, dev);
  unpack<49>(&r[0][0], R, dR);
  unpack<49>(&l[0][0], L, dL);
}

// kind 0: flux_roe (returns its HLLE-fallback flag), 1: flux_hlle
int ora_tlm_flux(double gamma, int kind, const double* ul, const double* dul, const double* ur,
                 const double* dur, const double* wl, const double* dwl, const double* wr,
                 const double* dwr, double bx, double dbx, double* f, double* df) {
  Cons<Dual> Ul, Ur;
  Prim<Dual> Wl, Wr;
  pack<7>(as_arr(Ul), ul, dul);
  pack<7>(as_arr(Ur), ur, dur);
  pack<7>(as_arr(Wl), wl, dwl);
  pack<7>(as_arr(Wr), wr, dwr);
  int used = 0;
  Cons<Dual> F = kind == 1 ? flux_hlle(Ul, Ur, Wl, Wr, Dual(bx, dbx), Eos(gamma))
                           : flux_roe(Ul, Ur, Wl, Wr, Dual(bx, dbx), Eos(gamma), &used);
  unpack<7>(as_arr(F), f, df);
  return used;
}

void ora_tlm_lr_states(double gamma, int n, const double* W, const double* dW, const double* Bxc,
                       const double* dBxc, double dtodx, int il, int iu, double* Wl, double* dWl,
                       double* Wr, double* dWr) {
  std::vector<Prim<Dual>> w(n), wl(n), wr(n);
  std::vector<Dual> b(n);
  for (int i = 0; i < n; i++) {
    pack<7>(as_arr(w[i]), W + 7 * i, dW + 7 * i);
    b[i] = Dual(Bxc[i], dBxc[i]);
  }
  lr_states(w.data(), b.data(), 0.0, dtodx, il, iu, wl.data(), wr.data(), Eos(gamma));
  for (int i = 0; i < n; i++) {
    unpack<7>(as_arr(wl[i]), Wl + 7 * i, dWl + 7 * i);
    unpack<7>(as_arr(wr[i]), Wr + 7 * i, dWr + 7 * i);
  }
}

// The rollout the adjoint differentiates: set_bvals, then nsteps times
// (integrate_2d with the fixed dt, set_bvals). state/dstate: 11 planes
// [Nx2+8][Nx1+8] in Grid order (d, M1, M2, M3, E, B1c, B2c, B3c, B1i, B2i, B3i).
void ora_tlm_rollout(int nx1, int nx2, double x1min, double x1max, double x2min, double x2max,
                     double gamma, const int* bc, int flux, double dt, int nsteps,
                     const double* state, const double* dstate, double* out, double* dout) {
  Grid<Dual> G;
  init_grid(G, nx1, nx2, x1min, x1max, x2min, x2max, gamma, 0.4);
  for (int k = 0; k < 4; k++) G.bc[k] = bc[k];
  G.flux = flux;
  const size_t np = (size_t)G.sx * G.sy;
  for (int q = 0; q < 11; q++) {
    std::vector<Dual>& p = *G.plane(q);
    for (size_t k = 0; k < np; k++) p[k] = Dual(state[q * np + k], dstate[q * np + k]);
  }
  Scratch<Dual> S;
  set_bvals(G);
  G.dt = dt;
  for (int s = 0; s < nsteps; s++) {
    integrate_2d(G, S);
    set_bvals(G);
  }
  for (int q = 0; q < 11; q++) {
    const std::vector<Dual>& p = *G.plane(q);
    for (size_t k = 0; k < np; k++) {
      out[q * np + k] = p[k].v;
      dout[q * np + k] = p[k].t;
    }
  }
}

// The same rollout with the branch log on: *digest and *count fingerprint
// every branch decision the rollout made on primal values (limiter and clamp
// selections, tracing signs, eigensystem degeneracies, upwinding and the HLLE
// fallback of flux_roe, corner-emf upwinding, |x|). Two states whose rollouts
// take the same branches give the same digest. The OpenMP thread count must
// be the same for the runs compared.
void ora_tlm_rollout_branches(int nx1, int nx2, double x1min, double x1max, double x2min,
                              double x2max, double gamma, const int* bc, int flux, double dt,
                              int nsteps, const double* state, const double* dstate,
                              double* out, double* dout, unsigned long long* digest,
                              unsigned long long* count) {
  BranchLog& L = branch_log();
  L.reset();
  L.on = true;
  ora_tlm_rollout(nx1, nx2, x1min, x1max, x2min, x2max, gamma, bc, flux, dt, nsteps, state,
                  dstate, out, dout);
  L.on = false;
  unsigned long long h = 0, n = 0;
  for (const BranchLog::Slot& sl : L.slot) {
    h += sl.h;
    n += sl.n;
  }
  *digest = h;
  *count = n;
}

// FP64 operations of nsteps calls of integrate_2d on a state (11 planes, as
// above; ghosts as given), counted on the reference algorithm as restated:
// ops[0] add/sub, ops[1] mul, ops[2] div, ops[3] sqrt. The max/min selections
// and comparisons are not counted.
void ora_count_ops(int nx1, int nx2, double x1min, double x1max, double x2min, double x2max,
                   double gamma, int flux, double dt, int nsteps, const double* state,
                   unsigned long long* ops) {
  Grid<Counted> G;
  init_grid(G, nx1, nx2, x1min, x1max, x2min, x2max, gamma, 0.4);
  G.flux = flux;
  const size_t np = (size_t)G.sx * G.sy;
  for (int q = 0; q < 11; q++) {
    std::vector<Counted>& p = *G.plane(q);
    for (size_t k = 0; k < np; k++) p[k] = Counted(state[q * np + k]);
  }
  Scratch<Counted> S;
  G.dt = dt;
  const int threads = omp_get_max_threads();
  omp_set_num_threads(1);  // the counters are per thread
  op_count() = OpCount();
  for (int s = 0; s < nsteps; s++) integrate_2d(G, S);
  omp_set_num_threads(threads);
  ops[0] = op_count().add;
  ops[1] = op_count().mul;
  ops[2] = op_count().div;
  ops[3] = op_count().sqrt;
}

}  // extern "C"
