// adjoint_kernels.hip -- reverse of one CTU+CT step (__integrate_2d__.c:100-537)
// followed by set_bvals (set_bvals.c:114-1031), for one instance.
//
// Given the adjoint of the step's output state (aout) and the forward scratch
// of the step recomputed from its input state S^n (predictor face states,
// conserved and primitive, and fluxes, predictor corner emf, corrector
// fluxes), the kernels below run the step's stages in reverse and accumulate
// the adjoint of S^n into ain:
//   fold        set_bvals: each ghost's adjoint is added to the cell it copies
//   update      Steps 9, 11a, 11b, 13 -> corrector fluxes, corner emf, identity
//   corner<C>   Step 9 corner emf -> corrector fluxes' By/Bz, emf3 at t^{n+1/2}
//   riemann_c   Steps 8b / 8c: corrector Riemann fluxes -> corrected face states
//   half        Step 7: emf3 at t^{n+1/2} -> half-step face fields, predictor fluxes
//   transverse  Steps 5, 6 and the half-step face fields of Step 4
//   corner<P>   Step 4 corner emf -> predictor fluxes' By/Bz, emf3 at t^n
//   emf_cc      Step 3
//   pred<X1/X2> Steps 1c-1e / 2c-2e: sources, prim->cons, predictor fluxes
//   plm_x1/x2   Steps 1b / 2b: the PLM reconstruction
//   plm_gather  Steps 1a / 2a: cons->prim
// Nonlinear stages recompute their forward values with the reference's
// evaluation order (this file is built with -ffp-contract=off, like the exact
// forward kernels), so every branch is the one the forward pass took. Linear
// stages scatter with FP64 atomic adds; a value reached through several
// stencil paths is summed in an unspecified order (roundoff-level).
#include <hip/hip_runtime.h>

#include "adjoint_device.hpp"
#include "adjoint_kernels.hpp"
#include "ghost_map.hpp"

namespace athmi {
namespace adjk {

using namespace adj;

__device__ __forceinline__ void ADD(double* plane, uint32_t k, double v) {
  unsafeAtomicAdd(plane + k, v);
}
__device__ __forceinline__ Cons7 ld7(const double* base, size_t ps, uint32_t k) {
  return Cons7{base[k], base[ps + k], base[2 * ps + k], base[3 * ps + k], base[4 * ps + k],
               base[5 * ps + k], base[6 * ps + k]};
}
__device__ __forceinline__ void st7(double* base, size_t ps, uint32_t k, const Cons7& u) {
  base[k] = u.d; base[ps + k] = u.mx; base[2 * ps + k] = u.my; base[3 * ps + k] = u.mz;
  base[4 * ps + k] = u.e; base[5 * ps + k] = u.by; base[6 * ps + k] = u.bz;
}
__device__ __forceinline__ void st7(double* base, size_t ps, uint32_t k, const Prim7& w) {
  base[k] = w.d; base[ps + k] = w.vx; base[2 * ps + k] = w.vy; base[3 * ps + k] = w.vz;
  base[4 * ps + k] = w.p; base[5 * ps + k] = w.by; base[6 * ps + k] = w.bz;
}
__device__ __forceinline__ Prim7 ld7p(const double* base, size_t ps, uint32_t k) {
  return Prim7{base[k], base[ps + k], base[2 * ps + k], base[3 * ps + k], base[4 * ps + k],
               base[5 * ps + k], base[6 * ps + k]};
}

// (j, i) of this thread in the range [j0, j1] x [i0, i1] tiled kTileX x kTileY
__device__ __forceinline__ bool range_ij(int j0, int j1, int i0, int i1, int& j, int& i) {
  i = i0 + (int)blockIdx.x * kTileX + (int)threadIdx.x;
  j = j0 + (int)blockIdx.y * kTileY + (int)threadIdx.y;
  return i <= i1 && j <= j1;
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kTileX* kTileY) ak_fold(AdjArgs a, int which) {
  const Inst& g = a.g;
  const int i = (int)blockIdx.x * kTileX + (int)threadIdx.x;
  const int j = (int)blockIdx.y * kTileY + (int)threadIdx.y;
  if (i >= g.sx || j >= g.sy) return;
  double* A = which ? a.ain : a.aout;
  const size_t ps = g.ps;
  const uint32_t k = (uint32_t)(j * g.pitch + i);
  for (int f = 0; f < NSTATE; f++) {
    int sj, si;
    bool neg;
    if (ghost_src(g, f, j, i, sj, si, neg)) {
      const double v = A[f * ps + k];
      if (v != 0.0) ADD(A + f * ps, (uint32_t)(sj * g.pitch + si), neg ? -v : v);
      A[f * ps + k] = 0.0;
    }
  }
}

// The same fold as an ordered gather (deterministic reverse): the thread of
// non-ghost cell (j, i) sums, field by field, the adjoints of every ghost
// whose fill copies that cell -- rows ascending, then columns -- and adds
// them once. The set_bvals map is separable: ghost row jj maps to a row
// independently of the column, and ghost column ii to a column, so the
// candidates are (rows mapping to j or j itself) x (columns mapping to i or i
// itself). Only cells near an edge can receive; the ghosts are cleared
// afterwards by ak_fold_clear.
__global__ void __launch_bounds__(kTileX* kTileY) ak_fold_gather(AdjArgs a) {
  const Inst& g = a.g;
  const int i = (int)blockIdx.x * kTileX + (int)threadIdx.x;
  const int j = (int)blockIdx.y * kTileY + (int)threadIdx.y;
  if (i >= g.sx || j >= g.sy) return;
  double* A = a.aout;
  const size_t ps = g.ps;
  const int pitch = g.pitch;
  for (int f = 0; f < NSTATE; f++) {
    const int ie_x = (f == F_B1I) ? g.ie + 1 : g.ie;
    const int je_y = (f == F_B2I) ? g.je + 1 : g.je;
    if (i < g.is || i > ie_x || j < g.js || j > je_y) continue;  // ghosts receive nothing
    // receivers lie within 4 cells of an edge of the cell-centred range
    // [js, je] x [is, ie] (face planes: one more cell on the far side)
    const bool band = (j - g.js < 5) || (je_y - j < 5) || (i - g.is < 5) || (ie_x - i < 5);
    if (!band) continue;
    // ascending: ghost rows below, the row itself, ghost rows above
    int rows[9], cols[9], nr = 0, nc = 0;
    int sj, si;
    bool neg;
    for (int jj = 0; jj < g.js; jj++)
      if (ghost_src(g, f, jj, g.is, sj, si, neg) && sj == j) rows[nr++] = jj;
    rows[nr++] = j;
    for (int jj = je_y + 1; jj < g.sy; jj++)
      if (ghost_src(g, f, jj, g.is, sj, si, neg) && sj == j) rows[nr++] = jj;
    for (int ii = 0; ii < g.is; ii++)
      if (ghost_src(g, f, g.js, ii, sj, si, neg) && si == i) cols[nc++] = ii;
    cols[nc++] = i;
    for (int ii = ie_x + 1; ii < g.sx; ii++)
      if (ghost_src(g, f, g.js, ii, sj, si, neg) && si == i) cols[nc++] = ii;
    double sum = 0.0;
    bool any = false;
    for (int r = 0; r < nr; r++)
      for (int c = 0; c < nc; c++) {
        const int jj = rows[r], ii = cols[c];
        if (jj == j && ii == i) continue;
        if (ghost_src(g, f, jj, ii, sj, si, neg) && sj == j && si == i) {
          const double v = A[f * ps + (uint32_t)(jj * pitch + ii)];
          sum += neg ? -v : v;
          any = true;
        }
      }
    if (any) A[f * ps + (uint32_t)(j * pitch + i)] += sum;
  }
}
__global__ void __launch_bounds__(kTileX* kTileY) ak_fold_clear(AdjArgs a) {
  const Inst& g = a.g;
  const int i = (int)blockIdx.x * kTileX + (int)threadIdx.x;
  const int j = (int)blockIdx.y * kTileY + (int)threadIdx.y;
  if (i >= g.sx || j >= g.sy) return;
  const uint32_t k = (uint32_t)(j * g.pitch + i);
  for (int f = 0; f < NSTATE; f++) {
    int sj, si;
    bool neg;
    if (ghost_src(g, f, j, i, sj, si, neg)) a.aout[f * g.ps + k] = 0.0;
  }
}

// ALL: every adjoint plane (start of a reverse pass); otherwise ain and the
// NACC accumulated planes (every step)
__global__ void __launch_bounds__(256) ak_zero(AdjArgs a, int all) {
  const size_t n_in = (size_t)NSTATE * a.g.ps;
  const size_t n_ab = (size_t)(all ? NADJ : NACC) * a.g.ps;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_in + n_ab;
       e += (size_t)gridDim.x * blockDim.x) {
    if (e < n_in) a.ain[e] = 0.0;
    else a.ab[e - n_in] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Steps 9, 11a, 11b, 13 (:473-534) over [js,je+1] x [is,ie+1]; the corrector
// flux adjoints are stored over [js-1,je+1] x [is-1,ie+1], the faces the
// corrector Riemann reverse reads (zero where no update reads the flux)
__global__ void __launch_bounds__(kTileX* kTileY) ak_update(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (!range_ij(g.js - 1, g.je + 1, g.is - 1, g.ie + 1, j, i)) return;
  const size_t ps = g.ps;
  const int pitch = g.pitch;
  const double dtodx1 = a.dt / g.dx1, dtodx2 = a.dt / g.dx2;
  const uint32_t k = (uint32_t)(j * pitch + i);
  const double* ao = a.aout;
  // adjoint of the updated value of cell field q at (jj, ii); 0 off the interior.
  // B3i = B3c after the update (Step 13), so B3c carries both adjoints.
  auto cb = [&](int q, int jj, int ii) -> double {
    if (jj < g.js || jj > g.je || ii < g.is || ii > g.ie) return 0.0;
    const uint32_t kk = (uint32_t)(jj * pitch + ii);
    double v = ao[q * ps + kk];
    if (q == F_B3C) v += ao[F_B3I * ps + kk];
    return v;
  };
  if (j >= g.js && j <= g.je && i >= g.is && i <= g.ie) {
    const int q_id[6] = {F_D, F_M1, F_M2, F_M3, F_E, F_B3C};
    for (int n = 0; n < 6; n++) a.ain[q_id[n] * ps + k] += cb(q_id[n], j, i);
  }
  // corrector fluxes; the B2c update of 11a and the B1c update of 11b are
  // overwritten by Step 13, so those components get no adjoint there
  if (i >= g.is) {
    const int q1[7] = {F_D, F_M1, F_M2, F_M3, F_E, -1, F_B3C};
    double* F = a.ab + A_F1C * ps;
    for (int n = 0; n < 7; n++)
      F[n * ps + k] = (q1[n] < 0) ? 0.0 : dtodx1 * (cb(q1[n], j, i) - cb(q1[n], j, i - 1));
  }
  if (j >= g.js) {
    const int q2[7] = {F_D, F_M2, F_M3, F_M1, F_E, F_B3C, -1};
    double* F = a.ab + A_F2C * ps;
    for (int n = 0; n < 7; n++)
      F[n * ps + k] = (q2[n] < 0) ? 0.0 : dtodx2 * (cb(q2[n], j, i) - cb(q2[n], j - 1, i));
  }
  if (j < g.js || i < g.is) return;
  // adjoint of the updated face fields including B1c, B2c = face averages
  auto b1t = [&](int jj, int ii) -> double {
    if (jj < g.js || jj > g.je || ii < g.is || ii > g.ie + 1) return 0.0;
    return ao[F_B1I * ps + (uint32_t)(jj * pitch + ii)] +
           0.5 * (cb(F_B1C, jj, ii) + cb(F_B1C, jj, ii - 1));
  };
  auto b2t = [&](int jj, int ii) -> double {
    if (jj < g.js || jj > g.je + 1 || ii < g.is || ii > g.ie) return 0.0;
    return ao[F_B2I * ps + (uint32_t)(jj * pitch + ii)] +
           0.5 * (cb(F_B2C, jj, ii) + cb(F_B2C, jj - 1, ii));
  };
  if (j <= g.je) a.ain[F_B1I * ps + k] += b1t(j, i);
  if (i <= g.ie) a.ain[F_B2I * ps + k] += b2t(j, i);
  a.ab[A_EMFC * ps + k] = dtodx2 * (b1t(j, i) - b1t(j - 1, i)) + dtodx1 * (b2t(j, i - 1) - b2t(j, i));
}

// ---------------------------------------------------------------------------
// integrate_emf3_corner (:632-702) reversed. PRED: Step 4 over
// [js-1,je+2]x[is-1,ie+2]; otherwise Step 9 over [js,je+1]x[is,ie+1].
__device__ __forceinline__ double upwind_weight(double mass_flux) {
  return (mass_flux > 0.0) ? 1.0 : ((mass_flux < 0.0) ? 0.0 : 0.5);
}

template <bool PRED>
__global__ void __launch_bounds__(kTileX* kTileY) ak_corner(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (PRED) {
    if (!range_ij(g.js - 1, g.je + 2, g.is - 1, g.ie + 2, j, i)) return;
  } else {
    if (!range_ij(g.js, g.je + 1, g.is, g.ie + 1, j, i)) return;
  }
  // deterministic reverse: one launch per (j, i) parity class; corners of one
  // class touch disjoint flux and emf entries, so no entry takes two adds
  if (a.det && (((j & 1) << 1) | (i & 1)) != a.color) return;
  const size_t ps = g.ps;
  const int pitch = g.pitch;
  const uint32_t k = (uint32_t)(j * pitch + i);
  const double q = 0.25 * a.ab[(PRED ? A_EMFP : A_EMFC) * ps + k];
  if (q == 0.0) return;
  const double* F1 = a.sc + (PRED ? S_F1 : S_F1C) * ps;  // mass fluxes pick the upwind side
  const double* F2 = a.sc + (PRED ? S_F2 : S_F2C) * ps;
  double* F1b = a.ab + (PRED ? A_F1 : A_F1C) * ps;
  double* F2b = a.ab + (PRED ? A_F2 : A_F2C) * ps;
  double* eb = a.ab + (PRED ? A_ECC : A_ECH) * ps;
  const uint32_t kjm = k - pitch, kim = k - 1, kjmim = k - pitch - 1;
  double c_f1jm = 0, c_f1 = 0, c_f2im = 0, c_f2 = 0, c_ejmim = 0, c_ejm = 0, c_eim = 0, c_e00 = 0;
  double w = upwind_weight(F1[kjm]);  // emf_l2
  c_f1jm -= 1.0; c_f2im += w; c_ejmim -= w; c_f2 += 1.0 - w; c_ejm -= 1.0 - w;
  w = upwind_weight(F1[k]);  // emf_r2
  c_f1 -= 1.0; c_f2im += w; c_eim -= w; c_f2 += 1.0 - w; c_e00 -= 1.0 - w;
  w = upwind_weight(F2[kim]);  // emf_l1
  c_f2im += 1.0; c_f1jm -= w; c_ejmim -= w; c_f1 -= 1.0 - w; c_eim -= 1.0 - w;
  w = upwind_weight(F2[k]);  // emf_r1
  c_f2 += 1.0; c_f1jm -= w; c_ejm -= w; c_f1 -= 1.0 - w; c_e00 -= 1.0 - w;
  ADD(F1b + 5 * ps, kjm, q * c_f1jm);
  ADD(F1b + 5 * ps, k, q * c_f1);
  ADD(F2b + 6 * ps, kim, q * c_f2im);
  ADD(F2b + 6 * ps, k, q * c_f2);
  ADD(eb, kjmim, q * c_ejmim);
  ADD(eb, kjm, q * c_ejm);
  ADD(eb, kim, q * c_eim);
  ADD(eb, k, q * c_e00);
}

// ---------------------------------------------------------------------------
// Forward values shared by the corrector stages: half-step face fields
// (:266-275) and the corrected face states (Steps 5, 6), in the exact forward
// order.
struct CorrCtx {
  const Inst& g;
  const double* st;
  const double* sc;
  const double* emf;  // predictor corner emf
  size_t ps;
  int pitch;
  double hdtodx1, hdtodx2;
  __device__ double b1h(uint32_t kk) const {
    return st[F_B1I * ps + kk] - hdtodx2 * (emf[kk + pitch] - emf[kk]);
  }
  __device__ double b2h(uint32_t kk) const {
    return st[F_B2I * ps + kk] + hdtodx1 * (emf[kk + 1] - emf[kk]);
  }
  __device__ double s(int f, uint32_t kk) const { return st[f * ps + kk]; }
  // x1 face k: Step 5a with x2 fluxes at (j+1,i-1),(j,i-1) / (j+1,i),(j,i) and 5b
  __device__ void states_x1(uint32_t k, Cons7& L, Cons7& R) const {
    const double* F2 = sc + S_F2 * ps;
    L = ld7(sc + S_UL1 * ps, ps, k);
    R = ld7(sc + S_UR1 * ps, ps, k);
    const uint32_t a1 = k + pitch - 1, a0 = k - 1, c1 = k + pitch, c0 = k;
    L.d = L.d - hdtodx2 * (F2[a1] - F2[a0]);
    L.mx = L.mx - hdtodx2 * (F2[3 * ps + a1] - F2[3 * ps + a0]);
    L.my = L.my - hdtodx2 * (F2[1 * ps + a1] - F2[1 * ps + a0]);
    L.mz = L.mz - hdtodx2 * (F2[2 * ps + a1] - F2[2 * ps + a0]);
    L.e = L.e - hdtodx2 * (F2[4 * ps + a1] - F2[4 * ps + a0]);
    L.bz = L.bz - hdtodx2 * (F2[5 * ps + a1] - F2[5 * ps + a0]);
    R.d = R.d - hdtodx2 * (F2[c1] - F2[c0]);
    R.mx = R.mx - hdtodx2 * (F2[3 * ps + c1] - F2[3 * ps + c0]);
    R.my = R.my - hdtodx2 * (F2[1 * ps + c1] - F2[1 * ps + c0]);
    R.mz = R.mz - hdtodx2 * (F2[2 * ps + c1] - F2[2 * ps + c0]);
    R.e = R.e - hdtodx2 * (F2[4 * ps + c1] - F2[4 * ps + c0]);
    R.bz = R.bz - hdtodx2 * (F2[5 * ps + c1] - F2[5 * ps + c0]);
    const double hdtodx1_ = hdtodx1;
    double dbx = s(F_B1I, k) - s(F_B1I, k - 1);
    double b1 = s(F_B1C, k - 1), b2 = s(F_B2C, k - 1), b3 = s(F_B3C, k - 1);
    double v3 = s(F_M3, k - 1) / s(F_D, k - 1);
    L.mx = L.mx + hdtodx1_ * b1 * dbx;
    L.my = L.my + hdtodx1_ * b2 * dbx;
    L.mz = L.mz + hdtodx1_ * b3 * dbx;
    L.bz = L.bz + hdtodx1_ * v3 * dbx;
    L.e = L.e + hdtodx1_ * b3 * v3 * dbx;
    dbx = s(F_B1I, k + 1) - s(F_B1I, k);
    b1 = s(F_B1C, k); b2 = s(F_B2C, k); b3 = s(F_B3C, k);
    v3 = s(F_M3, k) / s(F_D, k);
    R.mx = R.mx + hdtodx1_ * b1 * dbx;
    R.my = R.my + hdtodx1_ * b2 * dbx;
    R.mz = R.mz + hdtodx1_ * b3 * dbx;
    R.bz = R.bz + hdtodx1_ * v3 * dbx;
    R.e = R.e + hdtodx1_ * b3 * v3 * dbx;
  }
  // x2 face k: Step 6a with x1 fluxes at (j-1,i+1),(j-1,i) / (j,i+1),(j,i) and 6b
  __device__ void states_x2(uint32_t k, Cons7& L, Cons7& R) const {
    const double* F1 = sc + S_F1 * ps;
    L = ld7(sc + S_UL2 * ps, ps, k);
    R = ld7(sc + S_UR2 * ps, ps, k);
    const uint32_t a1 = k - pitch + 1, a0 = k - pitch, c1 = k + 1, c0 = k;
    L.d = L.d - hdtodx1 * (F1[a1] - F1[a0]);
    L.mx = L.mx - hdtodx1 * (F1[2 * ps + a1] - F1[2 * ps + a0]);
    L.my = L.my - hdtodx1 * (F1[3 * ps + a1] - F1[3 * ps + a0]);
    L.mz = L.mz - hdtodx1 * (F1[1 * ps + a1] - F1[1 * ps + a0]);
    L.e = L.e - hdtodx1 * (F1[4 * ps + a1] - F1[4 * ps + a0]);
    L.by = L.by - hdtodx1 * (F1[6 * ps + a1] - F1[6 * ps + a0]);
    R.d = R.d - hdtodx1 * (F1[c1] - F1[c0]);
    R.mx = R.mx - hdtodx1 * (F1[2 * ps + c1] - F1[2 * ps + c0]);
    R.my = R.my - hdtodx1 * (F1[3 * ps + c1] - F1[3 * ps + c0]);
    R.mz = R.mz - hdtodx1 * (F1[1 * ps + c1] - F1[1 * ps + c0]);
    R.e = R.e - hdtodx1 * (F1[4 * ps + c1] - F1[4 * ps + c0]);
    R.by = R.by - hdtodx1 * (F1[6 * ps + c1] - F1[6 * ps + c0]);
    const uint32_t km = k - pitch;
    double dby = s(F_B2I, k) - s(F_B2I, km);
    double b1 = s(F_B1C, km), b2 = s(F_B2C, km), b3 = s(F_B3C, km);
    double v3 = s(F_M3, km) / s(F_D, km);
    L.mz = L.mz + hdtodx2 * b1 * dby;
    L.mx = L.mx + hdtodx2 * b2 * dby;
    L.my = L.my + hdtodx2 * b3 * dby;
    L.by = L.by + hdtodx2 * v3 * dby;
    L.e = L.e + hdtodx2 * b3 * v3 * dby;
    dby = s(F_B2I, k + pitch) - s(F_B2I, k);
    b1 = s(F_B1C, k); b2 = s(F_B2C, k); b3 = s(F_B3C, k);
    v3 = s(F_M3, k) / s(F_D, k);
    R.mz = R.mz + hdtodx2 * b1 * dby;
    R.mx = R.mx + hdtodx2 * b2 * dby;
    R.my = R.my + hdtodx2 * b3 * dby;
    R.by = R.by + hdtodx2 * v3 * dby;
    R.e = R.e + hdtodx2 * b3 * v3 * dby;
  }
};

__device__ __forceinline__ CorrCtx corr_ctx(const AdjArgs& a) {
  const double dtodx1 = a.dt / a.g.dx1, dtodx2 = a.dt / a.g.dx2;
  return CorrCtx{a.g, a.st, a.sc, a.emfp, a.g.ps, a.g.pitch, 0.5 * dtodx1, 0.5 * dtodx2};
}

// Step 8b (X1: x1 faces [js-1,je+1] x [is,ie+1], :438-447) or 8c (x2 faces
// [js,je+1] x [is-1,ie+1], :455-464); the two are independent
template <bool X1, int FLUX>  // FLUX: the grid's Riemann solver
__global__ void __launch_bounds__(kTileX* kTileY) ak_riemann_c(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (X1) {
    if (!range_ij(g.js - 1, g.je + 1, g.is, g.ie + 1, j, i)) return;
  } else {
    if (!range_ij(g.js, g.je + 1, g.is - 1, g.ie + 1, j, i)) return;
  }
  const CorrCtx c = corr_ctx(a);
  const size_t ps = g.ps;
  const uint32_t k = (uint32_t)(j * g.pitch + i);
  const EosC eos{g.gamma, g.g1, g.g2};
  if (X1) {
    Cons7 L, R;
    c.states_x1(k, L, R);
    const double bx = c.b1h(k);
    const Prim7 wl = cons_to_prim(L, bx, eos), wr = cons_to_prim(R, bx, eos);
    Cons7 Lb = zero_cons(), Rb = zero_cons();
    Prim7 Wlb = zero_prim(), Wrb = zero_prim();
    double bxb = 0.0;
    riemann_b_of<FLUX>(L, R, wl, wr, bx, eos, ld7(a.ab + A_F1C * ps, ps, k), Lb, Rb, Wlb, Wrb, bxb);
    cons_to_prim_b(L, bx, eos, Wlb, Lb, bxb);
    cons_to_prim_b(R, bx, eos, Wrb, Rb, bxb);
    st7(a.ab + A_UL1 * ps, ps, k, Lb);
    st7(a.ab + A_UR1 * ps, ps, k, Rb);
    a.ab[A_B1H * ps + k] = bxb;
  } else {
    Cons7 L, R;
    c.states_x2(k, L, R);
    const double bx = c.b2h(k);
    const Prim7 wl = cons_to_prim(L, bx, eos), wr = cons_to_prim(R, bx, eos);
    Cons7 Lb = zero_cons(), Rb = zero_cons();
    Prim7 Wlb = zero_prim(), Wrb = zero_prim();
    double bxb = 0.0;
    riemann_b_of<FLUX>(L, R, wl, wr, bx, eos, ld7(a.ab + A_F2C * ps, ps, k), Lb, Rb, Wlb, Wrb, bxb);
    cons_to_prim_b(L, bx, eos, Wlb, Lb, bxb);
    cons_to_prim_b(R, bx, eos, Wrb, Rb, bxb);
    st7(a.ab + A_UL2 * ps, ps, k, Lb);
    st7(a.ab + A_UR2 * ps, ps, k, Rb);
    a.ab[A_B2H * ps + k] = bxb;
  }
}

// Step 7 (:402-431) reversed per cell of [js-1,je+1] x [is-1,ie+1]: stores
// the adjoints of the cell's half-step d, M1, M2 and B1c, B2c; the transverse
// reverse hands them on to the predictor fluxes and the half-step face fields.
__global__ void __launch_bounds__(kTileX* kTileY) ak_half(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (!range_ij(g.js - 1, g.je + 1, g.is - 1, g.ie + 1, j, i)) return;
  const size_t ps = g.ps;
  const int pitch = g.pitch;
  const uint32_t k = (uint32_t)(j * pitch + i);
  const double Eb = a.ab[A_ECH * ps + k];
  double dhb = 0.0, m1b = 0.0, m2b = 0.0, b1cb = 0.0, b2cb = 0.0;
  if (Eb != 0.0) {
    const CorrCtx c = corr_ctx(a);
    const double h1 = c.hdtodx1, h2 = c.hdtodx2;
    const double* F1 = a.sc + S_F1 * ps;
    const double* F2 = a.sc + S_F2 * ps;
    const double dhalf = c.s(F_D, k) - h1 * (F1[k + 1] - F1[k]) - h2 * (F2[k + pitch] - F2[k]);
    const double m1 = c.s(F_M1, k) - h1 * (F1[ps + k + 1] - F1[ps + k]) -
                      h2 * (F2[3 * ps + k + pitch] - F2[3 * ps + k]);
    const double m2 = c.s(F_M2, k) - h1 * (F1[2 * ps + k + 1] - F1[2 * ps + k]) -
                      h2 * (F2[ps + k + pitch] - F2[ps + k]);
    const double b1c = 0.5 * (c.b1h(k) + c.b1h(k + 1));
    const double b2c = 0.5 * (c.b2h(k) + c.b2h(k + pitch));
    const double emf = (b1c * m2 - b2c * m1) / dhalf;
    b1cb = adiv(Eb * m2, dhalf);
    m2b = adiv(Eb * b1c, dhalf);
    b2cb = -adiv(Eb * m1, dhalf);
    m1b = -adiv(Eb * b2c, dhalf);
    dhb = -adiv(Eb * emf, dhalf);
    a.ain[F_D * ps + k] += dhb;
    a.ain[F_M1 * ps + k] += m1b;
    a.ain[F_M2 * ps + k] += m2b;
  }
  double* ab = a.ab;
  ab[A_HD * ps + k] = dhb;
  ab[A_HM1 * ps + k] = m1b;
  ab[A_HM2 * ps + k] = m2b;
  ab[A_HB1 * ps + k] = b1cb;
  ab[A_HB2 * ps + k] = b2cb;
}

// Steps 5, 6 (:284-393), the half-step face fields of Step 4 (:263-275) and
// the flux part of Step 7, reversed as a gather: each thread owns one cell /
// face / corner position c of [js-2,je+2] x [is-2,ie+2] and stores the
// adjoints of the predictor fluxes and the predictor corner emf there, and
// adds the state adjoints of c. Every adjoint read is zero outside the range
// its writer covers, so no position needs a range test.
//
// The x1-face states at faces c+1 (left) and c (right) both take their 1/2 dt
// source terms from cell c and the face-field difference B1i[c+1] - B1i[c]
// (:296-330), and both are corrected with the x2 fluxes at c and c+pitch
// (left at c+1: at c+1-1): with X1(c) = Lb(c+1) + Rb(c) the sources' adjoint
// is linear in X1(c), and the x2 flux at c receives h2 (X1(c) - X1(c-pitch)).
// Likewise for x2 faces with X2(c) = Lb(c+pitch) + Rb(c).
__global__ void __launch_bounds__(kTileX* kTileY) ak_transverse(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (!range_ij(g.js - 2, g.je + 2, g.is - 2, g.ie + 2, j, i)) return;
  const size_t ps = g.ps;
  const uint32_t pitch = (uint32_t)g.pitch;
  const uint32_t c = (uint32_t)j * pitch + (uint32_t)i;
  const double h1 = 0.5 * a.dt / g.dx1, h2 = 0.5 * a.dt / g.dx2;
  const double* st = a.st;
  const double* ab = a.ab;
  double* ain = a.ain;
  // X1 and X2 components in Cons7 order (d, mx, my, mz, e, by, bz)
  auto X1 = [&](uint32_t kk, int n) {
    return ab[(A_UL1 + n) * ps + kk + 1] + ab[(A_UR1 + n) * ps + kk];
  };
  auto X2 = [&](uint32_t kk, int n) {
    return ab[(A_UL2 + n) * ps + kk + pitch] + ab[(A_UR2 + n) * ps + kk];
  };
  // h_dir * (the source terms' adjoint on the face field), per cell
  //   X.(b1, b2, b3, bt) components: x1 states (mx, my, mz, bz); x2 states
  //   (d, M2, M3, M1, E, B3, B1) -> (mz, mx, my, by)
  auto dbb = [&](uint32_t kk, double h, double xb1, double xb2, double xb3, double xbt,
                 double xe) {
    const double b1 = st[F_B1C * ps + kk], b2 = st[F_B2C * ps + kk], b3 = st[F_B3C * ps + kk];
    const double v3 = st[F_M3 * ps + kk] / st[F_D * ps + kk];
    return h * (xb1 * b1 + xb2 * b2 + xb3 * b3 + xbt * v3 + xe * b3 * v3);
  };
  auto dbb1 = [&](uint32_t kk) {
    return dbb(kk, h1, X1(kk, 1), X1(kk, 2), X1(kk, 3), X1(kk, 6), X1(kk, 4));
  };
  auto dbb2 = [&](uint32_t kk) {
    return dbb(kk, h2, X2(kk, 3), X2(kk, 1), X2(kk, 2), X2(kk, 5), X2(kk, 4));
  };
  // half-step face fields: Riemann part + Step 7's face averages
  auto hb1 = [&](uint32_t kk) {
    return ab[A_B1H * ps + kk] + 0.5 * (ab[A_HB1 * ps + kk] + ab[A_HB1 * ps + kk - 1]);
  };
  auto hb2 = [&](uint32_t kk) {
    return ab[A_B2H * ps + kk] + 0.5 * (ab[A_HB2 * ps + kk] + ab[A_HB2 * ps + kk - pitch]);
  };
  const double* HD = ab + A_HD * ps;
  const double* HM1 = ab + A_HM1 * ps;
  const double* HM2 = ab + A_HM2 * ps;
  // x2 predictor fluxes (d, M2, M3, M1, E, B3, B1): x1 states (d, mx, my, mz, e, bz) were
  // corrected with components (0, 3, 1, 2, 4, 5); Step 7 used d, M1, M2 = (0, 3, 1)
  {
    double* F2b = a.ab + A_F2 * ps;
    const int src[7] = {0, 2, 3, 1, 4, 6, -1};  // x1-state component feeding F2 component
    for (int n = 0; n < 7; n++) {
      double v = (src[n] < 0) ? 0.0 : h2 * (X1(c, src[n]) - X1(c - pitch, src[n]));
      if (n == 0) v += h2 * (HD[c] - HD[c - pitch]);
      if (n == 3) v += h2 * (HM1[c] - HM1[c - pitch]);
      if (n == 1) v += h2 * (HM2[c] - HM2[c - pitch]);
      F2b[n * ps + c] = v;
    }
  }
  // x1 predictor fluxes (d, M1, M2, M3, E, B2, B3): x2 states (d, mx, my, mz, e, by) were
  // corrected with components (0, 2, 3, 1, 4, 6); Step 7 used d, M1, M2 = (0, 1, 2)
  {
    double* F1b = a.ab + A_F1 * ps;
    const int src[7] = {0, 3, 1, 2, 4, -1, 5};  // x2-state component feeding F1 component
    for (int n = 0; n < 7; n++) {
      double v = (src[n] < 0) ? 0.0 : h1 * (X2(c, src[n]) - X2(c - 1, src[n]));
      if (n == 0) v += h1 * (HD[c] - HD[c - 1]);
      if (n == 1) v += h1 * (HM1[c] - HM1[c - 1]);
      if (n == 2) v += h1 * (HM2[c] - HM2[c - 1]);
      F1b[n * ps + c] = v;
    }
  }
  // predictor corner emf from the half-step face fields (:266-275)
  a.ab[A_EMFP * ps + c] = h2 * (hb1(c) - hb1(c - pitch)) + h1 * (hb2(c - 1) - hb2(c));
  // face fields: half-step fields and the source terms' face-field differences
  ain[F_B1I * ps + c] += hb1(c) + dbb1(c - 1) - dbb1(c);
  ain[F_B2I * ps + c] += hb2(c) + dbb2(c - pitch) - dbb2(c);
  // the source terms' cell-centred B and v3 = M3 / d of cell c
  {
    const double b3 = st[F_B3C * ps + c];
    const double d = st[F_D * ps + c], v3 = st[F_M3 * ps + c] / d;
    const double db1 = h1 * (st[F_B1I * ps + c + 1] - st[F_B1I * ps + c]);
    const double db2 = h2 * (st[F_B2I * ps + c + pitch] - st[F_B2I * ps + c]);
    const double x1e = X1(c, 4), x2e = X2(c, 4);
    ain[F_B1C * ps + c] += db1 * X1(c, 1) + db2 * X2(c, 3);
    ain[F_B2C * ps + c] += db1 * X1(c, 2) + db2 * X2(c, 1);
    ain[F_B3C * ps + c] += db1 * (X1(c, 3) + x1e * v3) + db2 * (X2(c, 2) + x2e * v3);
    const double v3b = db1 * (X1(c, 6) + x1e * b3) + db2 * (X2(c, 5) + x2e * b3);
    ain[F_M3 * ps + c] += adiv(v3b, d);
    ain[F_D * ps + c] += -adiv(v3b * v3, d);
  }
}

// Step 3 (:249-255) over [js-2,je+2] x [is-2,ie+2]
__global__ void __launch_bounds__(kTileX* kTileY) ak_emf_cc(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (!range_ij(g.js - 2, g.je + 2, g.is - 2, g.ie + 2, j, i)) return;
  const size_t ps = g.ps;
  const uint32_t k = (uint32_t)(j * g.pitch + i);
  const double Eb = a.ab[A_ECC * ps + k];
  if (Eb == 0.0) return;
  const double* st = a.st;
  const double d = st[F_D * ps + k], m1 = st[F_M1 * ps + k], m2 = st[F_M2 * ps + k];
  const double b1 = st[F_B1C * ps + k], b2 = st[F_B2C * ps + k];
  const double e = (b1 * m2 - b2 * m1) / d;
  ADD(a.ain + F_B1C * ps, k, adiv(Eb * m2, d));
  ADD(a.ain + F_M2 * ps, k, adiv(Eb * b1, d));
  ADD(a.ain + F_B2C * ps, k, -adiv(Eb * m1, d));
  ADD(a.ain + F_M1 * ps, k, -adiv(Eb * b2, d));
  ADD(a.ain + F_D * ps, k, -adiv(Eb * e, d));
}

// ---------------------------------------------------------------------------
// Steps 1c-1e / 2c-2e (:156-180, :216-241) reversed per predictor face: the
// Riemann flux, prim->cons and the 1/2 dt MHD source on the transverse field
// (By for x1 faces, Bz for x2 faces), from the face states the forward pass
// stored. X1: faces [js-2,je+2] x [is-1,ie+2]; x2: [js-1,je+2] x [is-2,ie+2].
// Writes the adjoints of the PLM outputs Wl, Wr (the source only adds to them).
template <bool X1, int FLUX>  // FLUX: the grid's Riemann solver
__global__ void __launch_bounds__(kTileX* kTileY) ak_pred(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (X1) {
    if (!range_ij(g.js - 2, g.je + 2, g.is - 1, g.ie + 2, j, i)) return;
  } else {
    if (!range_ij(g.js - 1, g.je + 2, g.is - 2, g.ie + 2, j, i)) return;
  }
  const size_t ps = g.ps;
  const uint32_t stride = X1 ? 1u : (uint32_t)g.pitch;
  const uint32_t k = (uint32_t)(j * g.pitch + i);
  const EosC eos{g.gamma, g.g1, g.g2};
  const double* st = a.st;
  const int bplane = X1 ? F_B1I : F_B2I;  // normal face field
  const int mplane = X1 ? F_M2 : F_M1;    // momentum of the source's velocity
  const double b0 = st[bplane * ps + k];
  Cons7 Ulb = ld7(a.ab + (X1 ? A_UL1 : A_UL2) * ps, ps, k);
  Cons7 Urb = ld7(a.ab + (X1 ? A_UR1 : A_UR2) * ps, ps, k);
  Prim7 wlb = zero_prim(), wrb = zero_prim();
  double b0b = 0.0;
  {
    const Prim7 wl = ld7p(a.pw + (X1 ? P_WL1 : P_WL2) * ps, ps, k);
    const Prim7 wr = ld7p(a.pw + (X1 ? P_WR1 : P_WR2) * ps, ps, k);
    riemann_b_of<FLUX>(ld7(a.sc + (X1 ? S_UL1 : S_UL2) * ps, ps, k),
              ld7(a.sc + (X1 ? S_UR1 : S_UR2) * ps, ps, k), wl, wr, b0, eos,
              ld7(a.ab + (X1 ? A_F1 : A_F2) * ps, ps, k), Ulb, Urb, wlb, wrb, b0b);
    prim_to_cons_b(wl, b0, eos, Ulb, wlb, b0b);
    prim_to_cons_b(wr, b0, eos, Urb, wrb, b0b);
  }
  // source: W.bt += 1/2 dt (M/d)[cell] (B[face+1] - B[face]) / dx, left state
  // from the cell behind the face, right state from the cell ahead
  const double dx = X1 ? g.dx1 : g.dx2;
  const double hdt = 0.5 * a.dt;
  const double bm = st[bplane * ps + k - stride], bp = st[bplane * ps + k + stride];
  const double dl = st[F_D * ps + k - stride], dr = st[F_D * ps + k];
  const double vl = st[mplane * ps + k - stride] / dl, vr = st[mplane * ps + k] / dr;
  const double slb = hdt * (X1 ? wlb.by : wlb.bz), srb = hdt * (X1 ? wrb.by : wrb.bz);
  const double dblb = adiv(slb * vl, dx), vlb = adiv(slb * (b0 - bm), dx);
  const double dbrb = adiv(srb * vr, dx), vrb = adiv(srb * (bp - b0), dx);
  if (a.det) {
    // stored per face; ak_plm_gather adds them to the cells in a fixed order
    double* pc = a.ab + (X1 ? A_PC1 : A_PC2) * ps;
    pc[k] = -dblb;
    pc[ps + k] = b0b + dblb - dbrb;
    pc[2 * ps + k] = dbrb;
    pc[3 * ps + k] = adiv(vlb, dl);
    pc[4 * ps + k] = -adiv(vlb * vl, dl);
    pc[5 * ps + k] = adiv(vrb, dr);
    pc[6 * ps + k] = -adiv(vrb * vr, dr);
  } else {
    ADD(a.ain + bplane * ps, k - stride, -dblb);
    ADD(a.ain + bplane * ps, k, b0b + dblb - dbrb);
    ADD(a.ain + bplane * ps, k + stride, dbrb);
    ADD(a.ain + mplane * ps, k - stride, adiv(vlb, dl));
    ADD(a.ain + F_D * ps, k - stride, -adiv(vlb * vl, dl));
    ADD(a.ain + mplane * ps, k, adiv(vrb, dr));
    ADD(a.ain + F_D * ps, k, -adiv(vrb * vr, dr));
  }
  st7(a.ab + (X1 ? A_WL1 : A_WL2) * ps, ps, k, wlb);
  st7(a.ab + (X1 ? A_WR1 : A_WR2) * ps, ps, k, wrb);
}

// Steps 1b / 2b reversed for one cell of [js-2,je+2] x [is-2,ie+2]: the PLM
// of cell k produced Wl at the next face and Wr at its own face. Stores the
// adjoints of the primitive states of cells k-1, k, k+1 and of Bxc[k] (22
// planes; zeros where the outputs' adjoints are zero); plm_gather hands them
// through cons->prim.
template <bool X1>
__global__ void __launch_bounds__(kTileX* kTileY) ak_plm(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (!range_ij(g.js - 2, g.je + 2, g.is - 2, g.ie + 2, j, i)) return;
  const size_t ps = g.ps;
  const uint32_t stride = X1 ? 1u : (uint32_t)g.pitch;
  const uint32_t k = (uint32_t)(j * g.pitch + i);
  double* dst = a.ab + (X1 ? A_PW1 : A_PW2) * ps;
  double olb[7], orb[7];
  bool any = false;
  {
    const Prim7 l = ld7p(a.ab + (X1 ? A_WL1 : A_WL2) * ps, ps, k + stride);
    const Prim7 r = ld7p(a.ab + (X1 ? A_WR1 : A_WR2) * ps, ps, k);
    for (int n = 0; n < 7; n++) {
      olb[n] = arr(l)[n];
      orb[n] = arr(r)[n];
      any = any || olb[n] != 0.0 || orb[n] != 0.0;
    }
  }
  if (!any) {
    for (int n = 0; n < 22; n++) dst[n * ps + k] = 0.0;
    return;
  }
  const EosC eos{g.gamma, g.g1, g.g2};
  const double* st = a.st;
  // cell state in the sweep's order and its normal field
  const int q[7] = {F_D, X1 ? F_M1 : F_M2, X1 ? F_M2 : F_M3, X1 ? F_M3 : F_M1, F_E,
                    X1 ? F_B2C : F_B3C, X1 ? F_B3C : F_B1C};
  const int qx = X1 ? F_B1C : F_B2C;
  auto cons_at = [&](uint32_t kk) {
    return Cons7{st[q[0] * ps + kk], st[q[1] * ps + kk], st[q[2] * ps + kk], st[q[3] * ps + kk],
                 st[q[4] * ps + kk], st[q[5] * ps + kk], st[q[6] * ps + kk]};
  };
  const uint32_t kc[3] = {k - stride, k, k + stride};
  Prim7 W[3];
  for (int c = 0; c < 3; c++) W[c] = cons_to_prim(cons_at(kc[c]), st[qx * ps + kc[c]], eos);
  // the stencil and eigensystem sit in LDS, one per thread, and each reverse
  // stage reads them again: the reverse then fits in registers without spills.
  // A slot is an odd number of doubles long, so that the 8-byte words one
  // field occupies in consecutive lanes' slots fall in distinct LDS banks
  // (68 doubles, an even stride, put lanes 4 or 8 to a bank)
  struct Slot {
    PlmCore c;
    double pad;
  };
  static_assert((sizeof(Slot) / 8) % 2 == 1, "odd slot stride in doubles");
  __shared__ Slot slot[kTileX * kTileY];
  PlmCore& P = slot[threadIdx.y * kTileX + threadIdx.x].c;
  P.prepare(W[0], W[1], W[2], st[qx * ps + k], a.dt / (X1 ? g.dx1 : g.dx2), eos);
  double wb[3][7];
  for (int c = 0; c < 3; c++)
    for (int n = 0; n < 7; n++) wb[c][n] = 0.0;
  double bxcb = 0.0;
  P.back(olb, orb, wb[0], wb[1], wb[2], bxcb);
  for (int c = 0; c < 3; c++)
    for (int n = 0; n < 7; n++) dst[(c * 7 + n) * ps + k] = wb[c][n];
  dst[21 * ps + k] = bxcb;
}

// Steps 1a / 2a reversed per cell: the primitive-state adjoint of cell c sums
// the PLM reverses of cells c+1 (its left neighbour slot), c and c-1, and goes
// through cons->prim once. X1: cells [js-2,je+2] x [is-3,ie+3]; x2:
// [js-3,je+3] x [is-2,ie+2]. Atomic adds: the two sweeps run concurrently.
template <bool X1>
__global__ void __launch_bounds__(kTileX* kTileY) ak_plm_gather(AdjArgs a) {
  const Inst& g = a.g;
  int j, i;
  if (X1) {
    if (!range_ij(g.js - 2, g.je + 2, g.is - 3, g.ie + 3, j, i)) return;
  } else {
    if (!range_ij(g.js - 3, g.je + 3, g.is - 2, g.ie + 2, j, i)) return;
  }
  const size_t ps = g.ps;
  const uint32_t stride = X1 ? 1u : (uint32_t)g.pitch;
  const uint32_t k = (uint32_t)(j * g.pitch + i);
  const double* src = a.ab + (X1 ? A_PW1 : A_PW2) * ps;
  double wb[7];
  bool any = false;
  for (int n = 0; n < 7; n++) {
    wb[n] = src[n * ps + k + stride] + src[(7 + n) * ps + k] + src[(14 + n) * ps + k - stride];
    any = any || wb[n] != 0.0;
  }
  double bxb = src[21 * ps + k];
  const int q[7] = {F_D, X1 ? F_M1 : F_M2, X1 ? F_M2 : F_M3, X1 ? F_M3 : F_M1, F_E,
                    X1 ? F_B2C : F_B3C, X1 ? F_B3C : F_B1C};
  const int qx = X1 ? F_B1C : F_B2C;
  if (a.det) {
    // the predictor-face source adjoints of faces k + stride, k and k - stride
    // (the deterministic reverse stores them per face instead of adding them)
    const double* pc = a.ab + (X1 ? A_PC1 : A_PC2) * ps;
    const int bplane = X1 ? F_B1I : F_B2I;
    const int mplane = X1 ? F_M2 : F_M1;
    const double bsum = pc[k + stride] + pc[ps + k] + pc[2 * ps + k - stride];
    const double msum = pc[3 * ps + k + stride] + pc[5 * ps + k];
    const double dsum = pc[4 * ps + k + stride] + pc[6 * ps + k];
    if (bsum != 0.0) ADD(a.ain + bplane * ps, k, bsum);
    if (msum != 0.0) ADD(a.ain + mplane * ps, k, msum);
    if (dsum != 0.0) ADD(a.ain + F_D * ps, k, dsum);
  }
  if (!any && bxb == 0.0) return;
  const EosC eos{g.gamma, g.g1, g.g2};
  const double* st = a.st;
  const Cons7 u{st[q[0] * ps + k], st[q[1] * ps + k], st[q[2] * ps + k], st[q[3] * ps + k],
                st[q[4] * ps + k], st[q[5] * ps + k], st[q[6] * ps + k]};
  Cons7 ub = zero_cons();
  const Prim7 Wb{wb[0], wb[1], wb[2], wb[3], wb[4], wb[5], wb[6]};
  cons_to_prim_b(u, st[qx * ps + k], eos, Wb, ub, bxb);
  for (int n = 0; n < 7; n++) ADD(a.ain + q[n] * ps, k, arr(ub)[n]);
  ADD(a.ain + qx * ps, k, bxb);
}

}  // namespace adjk
}  // namespace athmi

using namespace athmi;
using namespace athmi::adjk;

extern "C" hipError_t athmi_adjoint_launch(const AdjArgs* ap, int kernel, hipStream_t s) {
  const AdjArgs& a = *ap;
  const Inst& g = a.g;
  const dim3 b2(kTileX, kTileY);
  auto grid2 = [&](int j0, int j1, int i0, int i1) {
    return dim3((unsigned)cdiv(i1 - i0 + 1, kTileX), (unsigned)cdiv(j1 - j0 + 1, kTileY));
  };
  switch (kernel) {
    case AK_FOLD_OUT:
    case AK_FOLD_IN:
      hipLaunchKernelGGL(ak_fold, grid2(0, g.sy - 1, 0, g.sx - 1), b2, 0, s, a,
                         kernel == AK_FOLD_IN ? 1 : 0);
      break;
    case AK_ZERO:
    case AK_CLEAR:
      hipLaunchKernelGGL(ak_zero, dim3(2048), dim3(256), 0, s, a, kernel == AK_CLEAR ? 1 : 0);
      break;
    case AK_UPDATE:
      hipLaunchKernelGGL(ak_update, grid2(g.js - 1, g.je + 1, g.is - 1, g.ie + 1), b2, 0, s, a);
      break;
    case AK_CORNER_C:
      for (int c = 0; c < (a.det ? 4 : 1); c++) {
        AdjArgs ac = a;
        ac.color = c;
        hipLaunchKernelGGL(ak_corner<false>, grid2(g.js, g.je + 1, g.is, g.ie + 1), b2, 0, s, ac);
      }
      break;
    case AK_FOLD_OUT_DET:
      hipLaunchKernelGGL(ak_fold_gather, grid2(0, g.sy - 1, 0, g.sx - 1), b2, 0, s, a);
      hipLaunchKernelGGL(ak_fold_clear, grid2(0, g.sy - 1, 0, g.sx - 1), b2, 0, s, a);
      break;
    case AK_RIEMANN_C1:
      if (g.flux == kFluxHlle)
        hipLaunchKernelGGL((ak_riemann_c<true, kFluxHlle>), grid2(g.js - 1, g.je + 1, g.is, g.ie + 1),
                           b2, 0, s, a);
      else
        hipLaunchKernelGGL((ak_riemann_c<true, kFluxRoe>), grid2(g.js - 1, g.je + 1, g.is, g.ie + 1),
                           b2, 0, s, a);
      break;
    case AK_RIEMANN_C2:
      if (g.flux == kFluxHlle)
        hipLaunchKernelGGL((ak_riemann_c<false, kFluxHlle>), grid2(g.js, g.je + 1, g.is - 1, g.ie + 1),
                           b2, 0, s, a);
      else
        hipLaunchKernelGGL((ak_riemann_c<false, kFluxRoe>), grid2(g.js, g.je + 1, g.is - 1, g.ie + 1),
                           b2, 0, s, a);
      break;
    case AK_HALF:
      hipLaunchKernelGGL(ak_half, grid2(g.js - 1, g.je + 1, g.is - 1, g.ie + 1), b2, 0, s, a);
      break;
    case AK_TRANSVERSE:
      hipLaunchKernelGGL(ak_transverse, grid2(g.js - 2, g.je + 2, g.is - 2, g.ie + 2), b2, 0, s, a);
      break;
    case AK_CORNER_P:
      for (int c = 0; c < (a.det ? 4 : 1); c++) {
        AdjArgs ac = a;
        ac.color = c;
        hipLaunchKernelGGL(ak_corner<true>, grid2(g.js - 1, g.je + 2, g.is - 1, g.ie + 2), b2, 0, s,
                           ac);
      }
      break;
    case AK_EMF_CC:
      hipLaunchKernelGGL(ak_emf_cc, grid2(g.js - 2, g.je + 2, g.is - 2, g.ie + 2), b2, 0, s, a);
      break;
    case AK_PRED_X1:
      if (g.flux == kFluxHlle)
        hipLaunchKernelGGL((ak_pred<true, kFluxHlle>), grid2(g.js - 2, g.je + 2, g.is - 1, g.ie + 2),
                           b2, 0, s, a);
      else
        hipLaunchKernelGGL((ak_pred<true, kFluxRoe>), grid2(g.js - 2, g.je + 2, g.is - 1, g.ie + 2),
                           b2, 0, s, a);
      break;
    case AK_PLM_X1:
      hipLaunchKernelGGL(ak_plm<true>, grid2(g.js - 2, g.je + 2, g.is - 2, g.ie + 2), b2, 0, s, a);
      break;
    case AK_PRED_X2:
      if (g.flux == kFluxHlle)
        hipLaunchKernelGGL((ak_pred<false, kFluxHlle>), grid2(g.js - 1, g.je + 2, g.is - 2, g.ie + 2),
                           b2, 0, s, a);
      else
        hipLaunchKernelGGL((ak_pred<false, kFluxRoe>), grid2(g.js - 1, g.je + 2, g.is - 2, g.ie + 2),
                           b2, 0, s, a);
      break;
    case AK_PLM_X2:
      hipLaunchKernelGGL(ak_plm<false>, grid2(g.js - 2, g.je + 2, g.is - 2, g.ie + 2), b2, 0, s, a);
      break;
    case AK_PLMG_X1:
      hipLaunchKernelGGL(ak_plm_gather<true>, grid2(g.js - 2, g.je + 2, g.is - 3, g.ie + 3), b2, 0,
                         s, a);
      break;
    case AK_PLMG_X2:
      hipLaunchKernelGGL(ak_plm_gather<false>, grid2(g.js - 3, g.je + 3, g.is - 2, g.ie + 2), b2, 0,
                         s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
