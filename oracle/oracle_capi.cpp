// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY: C entry points over the CPU
// restatement (athena_oracle.hpp) for ctypes. Loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product path.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "athena_oracle.hpp"

using namespace ora;

namespace {
struct Handle {
  Grid<double> G;
  Scratch<double> S;
  std::vector<double> soln;  // cpaw2d initial state (8 planes)
  CpawParams cpaw{};
};
}  // namespace

extern "C" {

void* ora_grid_new(int nx1, int nx2, double x1min, double x1max, double x2min, double x2max,
                   double gamma, double cour_no, int ibc1, int obc1, int ibc2, int obc2, int flux) {
  Handle* h = new Handle();
  init_grid(h->G, nx1, nx2, x1min, x1max, x2min, x2max, gamma, cour_no);
  h->G.bc[0] = ibc1;
  h->G.bc[1] = obc1;
  h->G.bc[2] = ibc2;
  h->G.bc[3] = obc2;
  h->G.flux = flux;
  return h;
}

void ora_grid_free(void* p) { delete static_cast<Handle*>(p); }

void ora_info(void* p, int* iv, double* dv) {
  const Grid<double>& G = static_cast<Handle*>(p)->G;
  iv[0] = G.Nx1; iv[1] = G.Nx2; iv[2] = G.is; iv[3] = G.ie;
  iv[4] = G.js; iv[5] = G.je; iv[6] = G.nstep; iv[7] = G.idisp; iv[8] = G.jdisp;
  dv[0] = G.x1_0; dv[1] = G.x2_0; dv[2] = G.dx1; dv[3] = G.dx2;
  dv[4] = G.dt; dv[5] = G.time;
}

void ora_pgen_blast(void* p, double radius, double pamb, double prat, double b0, double angle) {
  pgen_blast(static_cast<Handle*>(p)->G, radius, pamb, prat, b0, angle);
}

void ora_pgen_blast_at(void* p, double radius, double pamb, double prat, double b0, double angle,
                       double cx, double cy) {
  pgen_blast(static_cast<Handle*>(p)->G, radius, pamb, prat, b0, angle, cx, cy);
}

void ora_pgen_field_loop(void* p, double rad, double amp, double vflow, int iprob) {
  pgen_field_loop(static_cast<Handle*>(p)->G, rad, amp, vflow, iprob);
}

double ora_pgen_cpaw2d(void* p, double angle, double b_par, double b_perp, double pres,
                       double v_par) {
  Handle* h = static_cast<Handle*>(p);
  h->cpaw = pgen_cpaw2d(h->G, angle, b_par, b_perp, pres, v_par, &h->soln);
  return h->cpaw.lambda;
}

// main.c:121-151 after the problem generator: ghost fill, then the first dt
void ora_start(void* p) {
  Handle* h = static_cast<Handle*>(p);
  set_bvals(h->G);
  new_dt(h->G);
}

// one iteration of main.c:221-318 in its ONLY_HOST form
int ora_step(void* p, double tlim) {
  Handle* h = static_cast<Handle*>(p);
  Grid<double>& G = h->G;
  if ((tlim - G.time) < G.dt) G.dt = (tlim - G.time);
  integrate_2d(G, h->S);
  G.nstep++;
  G.time += G.dt;
  new_dt(G);
  set_bvals(G);
  return 0;
}

int ora_run(void* p, double tlim, int nlim) {
  Handle* h = static_cast<Handle*>(p);
  int n = 0;
  while (h->G.time < tlim && (nlim < 0 || n < nlim)) {
    ora_step(p, tlim);
    n++;
  }
  return n;
}

void ora_integrate(void* p) {
  Handle* h = static_cast<Handle*>(p);
  integrate_2d(h->G, h->S);
}
void ora_new_dt(void* p) { new_dt(static_cast<Handle*>(p)->G); }
double ora_max_inverse_dt(void* p) { return max_inverse_dt(static_cast<Handle*>(p)->G); }
void ora_set_bvals(void* p) { set_bvals(static_cast<Handle*>(p)->G); }
void ora_set_time(void* p, double dt, double time, int nstep) {
  Grid<double>& G = static_cast<Handle*>(p)->G;
  G.dt = dt; G.time = time; G.nstep = nstep;
}

// U: AoS Gas[sy][sx] (8 doubles per cell, athena.h:43-52); B*: [sy][sx]
void ora_get(void* p, double* U, double* B1, double* B2, double* B3) {
  Grid<double>& G = static_cast<Handle*>(p)->G;
  const size_t n = (size_t)G.sx * G.sy;
  for (size_t k = 0; k < n; k++)
    for (int q = 0; q < 8; q++) U[8 * k + q] = (*G.plane(q))[k];
  std::memcpy(B1, G.B1i.data(), n * sizeof(double));
  std::memcpy(B2, G.B2i.data(), n * sizeof(double));
  std::memcpy(B3, G.B3i.data(), n * sizeof(double));
}

void ora_put(void* p, const double* U, const double* B1, const double* B2, const double* B3) {
  Grid<double>& G = static_cast<Handle*>(p)->G;
  const size_t n = (size_t)G.sx * G.sy;
  for (size_t k = 0; k < n; k++)
    for (int q = 0; q < 8; q++) (*G.plane(q))[k] = U[8 * k + q];
  std::memcpy(G.B1i.data(), B1, n * sizeof(double));
  std::memcpy(G.B2i.data(), B2, n * sizeof(double));
  std::memcpy(G.B3i.data(), B3, n * sizeof(double));
}

void ora_cpaw2d_l1(void* p, double* out) {
  Handle* h = static_cast<Handle*>(p);
  cpaw2d_l1(h->G, h->soln, out);
}

// ---- single routines ----------------------------------------------------
void ora_esys_prim(double gamma, const double* in, double* ev, double* rem, double* lem) {
  double R[7][7] = {}, L[7][7] = {};
  esys_prim<double>(in[0], in[1], in[2], in[3], in[4], in[5], Eos(gamma), ev, R, L);
  std::memcpy(rem, R, sizeof(R));
  std::memcpy(lem, L, sizeof(L));
}

void ora_esys_roe(double gamma, const double* in, double* ev, double* rem, double* lem,
                  int eig_only) {
  double R[7][7] = {}, L[7][7] = {};
  if (eig_only)
    esys_roe<double>(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9],
                     Eos(gamma), ev, nullptr, nullptr);
  else
    esys_roe<double>(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9],
                     Eos(gamma), ev, R, L);
  std::memcpy(rem, R, sizeof(R));
  std::memcpy(lem, L, sizeof(L));
}

static Cons<double> load_cons(const double* a) {
  return Cons<double>{a[0], a[1], a[2], a[3], a[4], a[5], a[6]};
}
static Prim<double> load_prim(const double* a) {
  return Prim<double>{a[0], a[1], a[2], a[3], a[4], a[5], a[6]};
}

int ora_flux_roe(double gamma, const double* ul, const double* ur, const double* wl,
                 const double* wr, double bxi, double* f) {
  int used = 0;
  Cons<double> F = flux_roe(load_cons(ul), load_cons(ur), load_prim(wl), load_prim(wr), bxi,
                            Eos(gamma), &used);
  std::memcpy(f, &F, sizeof(F));
  return used;
}

void ora_flux_hlle(double gamma, const double* ul, const double* ur, const double* wl,
                   const double* wr, double bxi, double* f) {
  Cons<double> F = flux_hlle(load_cons(ul), load_cons(ur), load_prim(wl), load_prim(wr), bxi,
                             Eos(gamma));
  std::memcpy(f, &F, sizeof(F));
}

void ora_cons_to_prim(double gamma, const double* u, double bx, double* w) {
  Prim<double> W = cons_to_prim(load_cons(u), bx, Eos(gamma));
  std::memcpy(w, &W, sizeof(W));
}

void ora_prim_to_cons(double gamma, const double* w, double bx, double* u) {
  Cons<double> U = prim_to_cons(load_prim(w), bx, Eos(gamma));
  std::memcpy(u, &U, sizeof(U));
}

void ora_lr_states(double gamma, int n, const double* W, const double* Bxc, double dt,
                   double dtodx, int il, int iu, double* Wl, double* Wr) {
  (void)n;
  lr_states(reinterpret_cast<const Prim<double>*>(W), Bxc, dt, dtodx, il, iu,
            reinterpret_cast<Prim<double>*>(Wl), reinterpret_cast<Prim<double>*>(Wr), Eos(gamma));
}

}  // extern "C"
