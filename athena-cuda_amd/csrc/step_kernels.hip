// step_kernels.hip -- the CTU+CT step (__integrate_2d__.c:100-537), new_dt
// (new_dt.c:25-103) and set_bvals (set_bvals.c:114-1031) as HIP kernels for
// gfx950 over a tile work-list spanning one or many grid instances.
//
// Pipeline of one step (all state in SoA planes, ping-pong buffers):
//   k_begin      clock: limits, tlim clamp of dt, reset of dt partials
//   k_pred_x1    Steps 1a-1e: x1 sweep, PLM via an LDS neighbour exchange
//   k_pred_x2    Steps 2a-2e: x2 sweep, each lane marches down a column
//   k_corner<P>  Step 4 corner emf (and Step 3 emf3_cc comes from k_pred_x1)
//   k_corr       Steps 4 (half-step face B), 5, 6, 7, 8b, 8c
//   k_corner<C>  Step 9 corner emf
//   k_final      Steps 9 (face B), 11a/11b, 13 and the new_dt max, fused
//   k_ghost      set_bvals on the new state
//   k_end        nstep++, time += dt, dt = min(2 dt, CourNo/max_dti), swap
//
// This file is compiled twice: ATHMI_VARIANT=exact with -ffp-contract=off
// (one rounding per operation, as the reference CPU build) and
// ATHMI_VARIANT=fast with FMA contraction.
#include <hip/hip_runtime.h>

#include "mhd_device.hpp"
#include "step_kernels.hpp"
#include "ghost_map.hpp"

#ifndef ATHMI_VARIANT
#error "ATHMI_VARIANT must name the arithmetic variant"
#endif

namespace athmi {
namespace ATHMI_VARIANT {

__device__ __forceinline__ int dcdiv(int a, int b) { return (a + b - 1) / b; }

// A value the same in every lane, moved to scalar registers (the compiler
// cannot always prove that values read through the instance descriptor are
// uniform, and keeps them in vector registers otherwise)
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uni(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <class T>
__device__ __forceinline__ T* uni(T* p) {
  const unsigned long long b = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(b & 0xffffffffu));
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}

// The hardware deals workgroups to the 8 XCDs round-robin (workgroup b to XCD
// b % 8). With remap, tile t of n is renumbered so that each XCD gets one
// contiguous run of tiles: row neighbours and the rows above and below then
// share that XCD's L2. Any bijection is correct; this one only changes
// which L2 serves which tile.
__device__ __forceinline__ int xcd_tile(int t, int n, int remap) {
  if (!remap) return t;
  const int x = t & 7, q = t >> 3, per = n >> 3, rem = n & 7;
  return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

// Position (tx, ty) of tile `local` in an ntx x nty grid of 2-D tiles. strip 0:
// row by row. strip s > 0: strips of s tile columns, each walked row by row
// before the next strip, so that the tile dispatched before a tile is mostly
// the one above it (s tiles back), whose rows -- this tile's upper halo -- the
// same XCD's L2 has just fetched. Any bijection is correct.
__device__ __forceinline__ void tile_pos(int local, int ntx, int nty, int strip, int& tx,
                                         int& ty) {
  if (strip <= 0 || strip >= ntx) {
    ty = local / ntx;
    tx = local - ty * ntx;
    return;
  }
  const int band = strip * nty;
  const int sidx = local / band;
  const int w = min(strip, ntx - sidx * strip);  // the last strip may be narrower
  const int r = local - sidx * band;
  ty = r / w;
  tx = sidx * strip + (r - ty * w);
}

// instance owning global tile t of kernel kind `kind` (tile0 ascending)
__device__ __forceinline__ int find_inst(const Inst* __restrict__ in, int n, int kind, int t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (in[mid].tile0[kind] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ Prim7 lds_prim(double (*s)[kX1Block], int t) {
  return Prim7{s[0][t], s[1][t], s[2][t], s[3][t], s[4][t], s[5][t], s[6][t]};
}
__device__ __forceinline__ void lds_put(double (*s)[kX1Block], int t, const Prim7& w) {
  s[0][t] = w.d; s[1][t] = w.vx; s[2][t] = w.vy; s[3][t] = w.vz;
  s[4][t] = w.p; s[5][t] = w.by; s[6][t] = w.bz;
}
// Global memory through a uniform plane base and a 32-bit element index: the
// hardware's saddr + voffset form, so all planes accessed at one cell share a
// single offset register (a plane holds < 2^29 doubles).
typedef __attribute__((address_space(1))) const char gconst_bytes;
typedef __attribute__((address_space(1))) char gbytes;
__device__ __forceinline__ double LD(const double* plane, uint32_t k) {
  return *(__attribute__((address_space(1))) const double*)((gconst_bytes*)plane + k * 8u);
}
__device__ __forceinline__ void ST(double* plane, uint32_t k, double v) {
  *(__attribute__((address_space(1))) double*)((gbytes*)plane + k * 8u) = v;
}

__device__ __forceinline__ Prim7 lds_prim64(double (*s)[kX2Block], int t) {
  return Prim7{s[0][t], s[1][t], s[2][t], s[3][t], s[4][t], s[5][t], s[6][t]};
}
__device__ __forceinline__ void lds_put64(double (*s)[kX2Block], int t, const Prim7& w) {
  s[0][t] = w.d; s[1][t] = w.vx; s[2][t] = w.vy; s[3][t] = w.vz;
  s[4][t] = w.p; s[5][t] = w.by; s[6][t] = w.bz;
}
// Keeps the compiler from forwarding earlier loads into a recomputation
// (emits no instruction)
__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }
// A phase boundary: memory accesses stay on their side (compiler_fence) and
// the scheduler moves no instruction across, so the live values of two
// phases are never interleaved (emits no instruction)
__device__ __forceinline__ void phase_fence() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

__device__ __forceinline__ void put7(double* base, size_t ps, uint32_t k, const Cons7& u) {
  ST(base + 0 * ps, k, u.d); ST(base + 1 * ps, k, u.mx); ST(base + 2 * ps, k, u.my);
  ST(base + 3 * ps, k, u.mz); ST(base + 4 * ps, k, u.e); ST(base + 5 * ps, k, u.by);
  ST(base + 6 * ps, k, u.bz);
}
__device__ __forceinline__ void put7(double* base, size_t ps, uint32_t k, const Prim7& w) {
  ST(base + 0 * ps, k, w.d); ST(base + 1 * ps, k, w.vx); ST(base + 2 * ps, k, w.vy);
  ST(base + 3 * ps, k, w.vz); ST(base + 4 * ps, k, w.p); ST(base + 5 * ps, k, w.by);
  ST(base + 6 * ps, k, w.bz);
}
__device__ __forceinline__ Cons7 get7(const double* base, size_t ps, uint32_t k) {
  return Cons7{LD(base + 0 * ps, k), LD(base + 1 * ps, k), LD(base + 2 * ps, k), LD(base + 3 * ps, k),
               LD(base + 4 * ps, k), LD(base + 5 * ps, k), LD(base + 6 * ps, k)};
}

__device__ __forceinline__ double block_max(double v, double* red) {
  // wave-level max with cross-lane shuffles, then across the block's waves
  // via LDS (red: one slot per wave)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double m = red[0];
  const int nw = (blockDim.x * blockDim.y) >> 6;
  for (int w = 1; w < nw; w++) m = dmax(m, red[w]);
  __syncthreads();
  return m;
}

__device__ __forceinline__ void atomic_max_pos(unsigned long long* slot, double v) {
  atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------
// Steps 1a-1e (__integrate_2d__.c:128-180): one row segment per tile. Each
// thread converts one cell to primitives; neighbours are exchanged in LDS for
// the characteristic PLM (lr_states_plm.c:95-245); each face then gets the
// 1/2 dt MHD source on By, prim->cons and the Riemann flux.
// LDS slots of one thread for flux_roe_reload's parked values (SoA over the
// workgroup, conflict-free); the corrector uses them to stay within the
// registers of two waves per SIMD
struct LdsPark {
  double* base;  // [14][nthreads]
  int stride, lane;
  __device__ void put(int q, double x) const { base[q * stride + lane] = x; }
  __device__ double get(int q) const { return base[q * stride + lane]; }
  __device__ void fence() const { compiler_fence(); }
};

// The x1 PLM's stencil (cells tid-1, tid, tid+1 of the row in LDS), read
// again after each fence so that it holds no registers in between
struct RowStencil {
  double (*s)[kX1Block];
  int t;
  __device__ void get(Prim7& m, Prim7& z, Prim7& p) const {
    m = lds_prim(s, t - 1); z = lds_prim(s, t); p = lds_prim(s, t + 1);
  }
  __device__ void fence() const { compiler_fence(); }
};

// SAVE: also store the primitive face states (Inst::prims). FLUX: the Riemann
// solver; a launch takes only the tiles of instances that use it, and with one
// solver per kernel the predictor fits two waves per SIMD.
template <bool SAVE, int FLUX>
__global__ void __launch_bounds__(kX1Block, 2) k_pred_x1(const Inst* __restrict__ insts,
                                                         int ninst, int ntiles, int remap) {
  __shared__ double sw[7][kX1Block];
  __shared__ double swl[7][kX1Block];
  __shared__ double swr[7][kX1Block];
  const int tid = threadIdx.x;
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_PRED_X1, t)];
    const Clock* clk = g.clk;
    if (!clk->active || g.flux != FLUX) continue;
    const int local = uni(t - g.tile0[K_PRED_X1]);
    const int gis = uni(g.is), gie = uni(g.ie);
    const int nseg = dcdiv(uni(g.nx1) + 4, kX1Faces);
    const int row = local / nseg, seg = local - row * nseg;
    const int j = uni(g.js) - 2 + row;
    const int f0 = gis - 1 + seg * kX1Faces;
    const int cell = f0 - 2 + tid;
    const double dt = uni(clk->dt);
    const double dx1 = uni(g.dx1);
    const double dtodx1 = dt / dx1;
    const double hdt = 0.5 * dt;
    const EosC eos{uni(g.gamma), uni(g.g1), uni(g.g2)};
    const double* __restrict__ st = uni(g.state[clk->cur]);
    double* __restrict__ sc = uni(g.scratch);
    double* __restrict__ pw = uni(g.prims);
    const size_t ps = (size_t)uni((int)g.ps);
    const uint32_t rk = (uint32_t)(j * uni(g.pitch));
    double bxc = 0.0;
    if (cell >= gis - 3 && cell <= gie + 3) {
      const uint32_t k = rk + (uint32_t)cell;
      const double d = LD(st + F_D * ps, k), m1 = LD(st + F_M1 * ps, k), m2 = LD(st + F_M2 * ps, k);
      const double b1c = LD(st + F_B1C * ps, k), b2c = LD(st + F_B2C * ps, k);
      const Cons7 u{d, m1, m2, LD(st + F_M3 * ps, k), LD(st + F_E * ps, k), b2c, LD(st + F_B3C * ps, k)};
      bxc = b1c;
      lds_put(sw, tid, cons_to_prim(u, bxc, eos));
      // Step 3 (:249-255): emf3 at t^n over [jl,ju] x [il,iu]
      if (cell >= gis - 2 && cell <= gie + 2 && tid >= 1 && tid <= kX1Block - 2)
        ST(sc + S_ECC * ps, k, (b1c * m2 - b2c * m1) / d);
    }
    __syncthreads();
    if (tid >= 1 && tid <= kX1Block - 2 && cell >= gis - 2 && cell <= gie + 2) {
      Prim7 wl, wr;
      plm_cell_s(RowStencil{sw, tid}, bxc, dtodx1, eos, wl, wr);
      lds_put(swl, tid + 1, wl);
      lds_put(swr, tid, wr);
    }
    __syncthreads();
    if (tid >= 2 && tid <= kX1Block - 2 && cell <= gie + 2) {
      const uint32_t k = rk + (uint32_t)cell;
      // face states: PLM output + MHD source terms for 1/2 dt (:156-164) +
      // prim->cons (:175-176); rebuilt from LDS and memory for the HLLE fallback
      auto build = [&](Cons7& ul, Cons7& ur, Prim7& wl, Prim7& wr, double& b0) {
        compiler_fence();
        wl = lds_prim(swl, tid);
        wr = lds_prim(swr, tid);
        const double bm = LD(st + F_B1I * ps, k - 1), bp = LD(st + F_B1I * ps, k + 1);
        b0 = LD(st + F_B1I * ps, k);
        double src = (LD(st + F_M2 * ps, k - 1) / LD(st + F_D * ps, k - 1)) * (b0 - bm) / dx1;
        wl.by = wl.by + hdt * src;
        src = (LD(st + F_M2 * ps, k) / LD(st + F_D * ps, k)) * (bp - b0) / dx1;
        wr.by = wr.by + hdt * src;
        ul = prim_to_cons(wl, b0, eos);
        ur = prim_to_cons(wr, b0, eos);
      };
      Cons7 ul, ur;
      Prim7 wl, wr;
      double b0;
      build(ul, ur, wl, wr, b0);
      if (SAVE) {
        put7(pw + P_WL1 * ps, ps, k, wl);
        put7(pw + P_WR1 * ps, ps, k, wr);
      }
      put7(sc + S_UL1 * ps, ps, k, ul);
      put7(sc + S_UR1 * ps, ps, k, ur);
      put7(sc + S_F1 * ps, ps, k, riemann_of<FLUX>(ul, ur, wl, wr, b0, eos, build));
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Steps 2a-2e (__integrate_2d__.c:189-241): the x2 sweep in the permuted
// order (d, M2, M3, M1, E, B3c, B1c) with Bx = B2c. Lanes own adjacent
// columns (coalesced rows); each lane marches down a segment of faces keeping
// the three-cell PLM stencil and the pending left state in registers.
// (SAVE and FLUX as for the x1 predictor; the marching stencil keeps this one
// at one wave per SIMD)
template <bool SAVE, int FLUX>
__global__ void __launch_bounds__(kX2Block) k_pred_x2(const Inst* __restrict__ insts, int ninst,
                                                      int ntiles, int seg_len, int remap) {
  __shared__ double sxl[7][kX2Block];
  __shared__ double sxr[7][kX2Block];
  const int lane = threadIdx.x;
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_PRED_X2, t)];
    const Clock* clk = g.clk;
    if (!clk->active || g.flux != FLUX) continue;
    const int local = t - g.tile0[K_PRED_X2];
    // chunks from column 0: each row segment starts on a cache line
    const int ncb = dcdiv(g.nx1 + 6, kX2Block);
    const int sg = local / ncb, cb = local - sg * ncb;
    const int i = cb * kX2Block + lane;
    if (i < g.is - 2) continue;
    if (i > g.ie + 2) continue;
    const int jf0 = g.js - 1 + sg * seg_len;
    const int jf1 = min(jf0 + seg_len - 1, g.je + 2);
    const double dt = clk->dt;
    const double dtodx2 = dt / g.dx2;
    const double hdt = 0.5 * dt;
    const EosC eos{g.gamma, g.g1, g.g2};
    const double* __restrict__ st = g.state[clk->cur];
    double* __restrict__ sc = g.scratch;
    const size_t ps = g.ps;
    const int pitch = g.pitch;
    auto prim_at = [&](int jj, double& bx) -> Prim7 {
      const uint32_t k = (uint32_t)(jj * pitch + i);
      const Cons7 u{LD(st + F_D * ps, k), LD(st + F_M2 * ps, k), LD(st + F_M3 * ps, k), LD(st + F_M1 * ps, k),
                    LD(st + F_E * ps, k), LD(st + F_B3C * ps, k), LD(st + F_B1C * ps, k)};
      bx = LD(st + F_B2C * ps, k);
      return cons_to_prim(u, bx, eos);
    };
    double bxa, bxb, bxc;
    Prim7 wa = prim_at(jf0 - 2, bxa);
    Prim7 wb = prim_at(jf0 - 1, bxb);
    Prim7 wc = prim_at(jf0, bxc);
    Prim7 wl_cur, wr_unused;
    plm_cell(wa, wb, wc, bxb, dtodx2, eos, wl_cur, wr_unused);
    // the conserved state of the row after next is loaded one iteration
    // ahead, so its latency hides behind the current face's work (the last
    // iteration reads at most row je+4, the outermost ghost row)
    auto raw_at = [&](int jj, Cons7& u, double& bx) {
      const uint32_t k = (uint32_t)(jj * pitch + i);
      u = Cons7{LD(st + F_D * ps, k), LD(st + F_M2 * ps, k), LD(st + F_M3 * ps, k), LD(st + F_M1 * ps, k),
                LD(st + F_E * ps, k), LD(st + F_B3C * ps, k), LD(st + F_B1C * ps, k)};
      bx = LD(st + F_B2C * ps, k);
    };
    Cons7 un;
    double bxn;
    raw_at(jf0 + 1, un, bxn);
    for (int j = jf0; j <= jf1; j++) {
      wa = wb;
      wb = wc;
      bxb = bxc;
      wc = cons_to_prim(un, bxn, eos);
      bxc = bxn;
      raw_at(j + 2, un, bxn);
      Prim7 wl_next, wr_cur;
      plm_cell(wa, wb, wc, bxb, dtodx2, eos, wl_next, wr_cur);
      const uint32_t k = (uint32_t)(j * pitch + i);
      lds_put64(sxl, lane, wl_cur);
      lds_put64(sxr, lane, wr_cur);
      // face states: PLM output + MHD source terms for 1/2 dt (:216-224) +
      // prim->cons (:235-236); rebuilt from LDS and memory for the HLLE fallback
      auto build = [&](Cons7& ul, Cons7& ur, Prim7& wl, Prim7& wr, double& b0) {
        compiler_fence();
        wl = lds_prim64(sxl, lane);
        wr = lds_prim64(sxr, lane);
        const double bm = LD(st + F_B2I * ps, k - pitch), bp = LD(st + F_B2I * ps, k + pitch);
        b0 = LD(st + F_B2I * ps, k);
        double src = (LD(st + F_M1 * ps, k - pitch) / LD(st + F_D * ps, k - pitch)) * (b0 - bm) / g.dx2;
        wl.bz = wl.bz + hdt * src;
        src = (LD(st + F_M1 * ps, k) / LD(st + F_D * ps, k)) * (bp - b0) / g.dx2;
        wr.bz = wr.bz + hdt * src;
        ul = prim_to_cons(wl, b0, eos);
        ur = prim_to_cons(wr, b0, eos);
      };
      Cons7 ul, ur;
      Prim7 wl, wr;
      double b0;
      build(ul, ur, wl, wr, b0);
      if (SAVE) {
        put7(g.prims + P_WL2 * ps, ps, k, wl);
        put7(g.prims + P_WR2 * ps, ps, k, wr);
      }
      put7(sc + S_UL2 * ps, ps, k, ul);
      put7(sc + S_UR2 * ps, ps, k, ur);
      put7(sc + S_F2 * ps, ps, k, riemann_of<FLUX>(ul, ur, wl, wr, b0, eos, build));
      wl_cur = wl_next;
    }
  }
}

// ---------------------------------------------------------------------------
// Steps 2a-2e again, for two waves per SIMD: the same march down a segment of
// x2 faces, but nothing of the stencil stays in registers across the Riemann
// flux. Each lane keeps five 7-double slots in LDS: the primitive states of
// rows j-1, j (and j+1 while the PLM of cell j runs), the pending left state
// Wl(j) and the right state Wr(j). Per face: convert row j+1 into the free
// slot, PLM of cell j (its Wl(j+1) overwrites the dead row j-1), then the face
// state build and the flux reading Wl(j), Wr(j) back from LDS. The slots
// rotate: row j-1's slot becomes the next pending Wl, Wl(j)'s the next free.
// The face-normal field and the 1/2 dt source term come from LDS as well
// (five doubles per lane): B2i of rows j and j+1, M1/d of the newest row,
// and the source s(r) = (M1/d)(r) (B2i(r+1) - B2i(r)) / dx2 of rows j-1 and
// j -- the one product both faces of cell r add (to Wr(r) and Wl(r+1)), so
// each row's M1/d and B2i are read and divided once.
// LDS: (5 x 7 + 5) x 64 doubles = 20 KiB per wave: eight waves fill a CU's
// 160 KiB.
constexpr int kX2Slots = 5;  // 4 rotating + the Wr slot
constexpr int kX2Aux = 5;    // B2i ring (2), s ring (2), M1/d carry
template <bool SAVE, int FLUX>
__global__ void __launch_bounds__(kX2Block, 2) k_pred_x2l(const Inst* __restrict__ insts,
                                                          int ninst, int ntiles, int seg_len,
                                                          int remap) {
  __shared__ double slots[kX2Slots][7][kX2Block];
  __shared__ double aux[kX2Aux][kX2Block];
  const int lane = threadIdx.x;
  // aux rows: B2i of row r at r & 1, s(r) at 2 + (r & 1), M1/d of the newest row
  auto b2i_slot = [&](int r) -> double& { return aux[r & 1][lane]; };
  auto src_slot = [&](int r) -> double& { return aux[2 + (r & 1)][lane]; };
  double& v1_carry = aux[4][lane];
  auto put = [&](int s, const Prim7& w) {
    slots[s][0][lane] = w.d; slots[s][1][lane] = w.vx; slots[s][2][lane] = w.vy;
    slots[s][3][lane] = w.vz; slots[s][4][lane] = w.p; slots[s][5][lane] = w.by;
    slots[s][6][lane] = w.bz;
  };
  auto get = [&](int s) -> Prim7 {
    return Prim7{slots[s][0][lane], slots[s][1][lane], slots[s][2][lane], slots[s][3][lane],
                 slots[s][4][lane], slots[s][5][lane], slots[s][6][lane]};
  };
  constexpr int kWr = 4;
  // the PLM's stencil, read from the slots again after each fence
  typedef decltype(get) Getter;
  struct LdsStencil {
    const Getter* rd;
    int a, b, c;
    __device__ void get(Prim7& m, Prim7& z, Prim7& p) const {
      m = (*rd)(a); z = (*rd)(b); p = (*rd)(c);
    }
    __device__ void fence() const { compiler_fence(); }
  };
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_PRED_X2, t)];
    const Clock* clk = g.clk;
    if (!clk->active || g.flux != FLUX) continue;
    const int local = uni(t - g.tile0[K_PRED_X2]);
    // chunks from column 0: each row segment starts on a cache line
    const int ncb = dcdiv(uni(g.nx1) + 6, kX2Block);
    const int sg = local / ncb, cb = local - sg * ncb;
    const int i = cb * kX2Block + lane;
    if (i < uni(g.is) - 2) continue;
    if (i > uni(g.ie) + 2) continue;
    const int jf0 = uni(g.js) - 1 + sg * seg_len;
    const int jf1 = min(jf0 + seg_len - 1, uni(g.je) + 2);
    const double dt = uni(clk->dt);
    const double dx2 = uni(g.dx2);
    const double dtodx2 = dt / dx2;
    const double hdt = 0.5 * dt;
    const EosC eos{uni(g.gamma), uni(g.g1), uni(g.g2)};
    const double* __restrict__ st = uni(g.state[clk->cur]);
    double* __restrict__ sc = uni(g.scratch);
    double* __restrict__ pw = uni(g.prims);
    const size_t ps = (size_t)uni((int)g.ps);
    const int pitch = uni(g.pitch);
    // cons->prim of row jj in the permuted order (d, M2, M3, M1, E, B3c, B1c), Bx = B2c;
    // with M1/d (the source term's division form) and B2i of the row
    auto load_row = [&](int jj, int slot, double& v1, double& b2i) {
      const uint32_t k = (uint32_t)(jj * pitch + i);
      const double d = LD(st + F_D * ps, k), m1 = LD(st + F_M1 * ps, k);
      const Cons7 u{d, LD(st + F_M2 * ps, k), LD(st + F_M3 * ps, k), m1,
                    LD(st + F_E * ps, k), LD(st + F_B3C * ps, k), LD(st + F_B1C * ps, k)};
      put(slot, cons_to_prim(u, LD(st + F_B2C * ps, k), eos));
      v1 = m1 / d;
      b2i = LD(st + F_B2I * ps, k);
    };
    auto bx_at = [&](int jj) { return LD(st + F_B2C * ps, (uint32_t)(jj * pitch + i)); };
    // start of the march: rows jf0-2, jf0-1, jf0, the source s(jf0-1) and
    // the PLM of cell jf0-1
    {
      double v1m, b2m, v10, b20, v1u, b2u;
      load_row(jf0 - 2, 0, v1u, b2u);
      load_row(jf0 - 1, 1, v1m, b2m);
      load_row(jf0, 2, v10, b20);
      src_slot(jf0 - 1) = v1m * (b20 - b2m) / dx2;
      b2i_slot(jf0) = b20;
      v1_carry = v10;
    }
    {
      Prim7 wl, wr_unused;
      plm_cell_s(LdsStencil{&get, 0, 1, 2}, bx_at(jf0 - 1), dtodx2, eos, wl, wr_unused);
      compiler_fence();
      put(0, wl);
    }
    int sP = 1, sQ = 2, sW = 0, sF = 3;  // rows j-1, j; pending Wl(j); free
    for (int j = jf0; j <= jf1; j++) {
      phase_fence();
      {
        // row j+1; then s(j) = (M1/d)(j) (B2i(j+1) - B2i(j)) / dx2 (:216-224)
        double v1n, b2n;
        load_row(j + 1, sF, v1n, b2n);
        src_slot(j) = v1_carry * (b2n - b2i_slot(j)) / dx2;
        b2i_slot(j + 1) = b2n;
        v1_carry = v1n;
      }
      phase_fence();
      {
        Prim7 wl_next, wr_cur;
        const LdsStencil sten{&get, sP, sQ, sF};
        plm_cell_s(sten, bx_at(j), dtodx2, eos, wl_next, wr_cur);
        compiler_fence();
        put(kWr, wr_cur);
        put(sP, wl_next);
      }
      phase_fence();
      const uint32_t k = (uint32_t)(j * pitch + i);
      const int sw = sW;
      // face states: PLM output + MHD source terms for 1/2 dt (:216-224) +
      // prim->cons (:235-236); rebuilt from LDS for the HLLE fallback
      auto build = [&](Cons7& ul, Cons7& ur, Prim7& wl, Prim7& wr, double& b0) {
        compiler_fence();
        wl = get(sw);
        wr = get(kWr);
        b0 = b2i_slot(j);
        wl.bz = wl.bz + hdt * src_slot(j - 1);
        wr.bz = wr.bz + hdt * src_slot(j);
        ul = prim_to_cons(wl, b0, eos);
        ur = prim_to_cons(wr, b0, eos);
      };
      Cons7 ul, ur;
      Prim7 wl, wr;
      double b0;
      build(ul, ur, wl, wr, b0);
      if (SAVE) {
        put7(pw + P_WL2 * ps, ps, k, wl);
        put7(pw + P_WR2 * ps, ps, k, wr);
      }
      put7(sc + S_UL2 * ps, ps, k, ul);
      put7(sc + S_UR2 * ps, ps, k, ur);
      put7(sc + S_F2 * ps, ps, k, riemann_of<FLUX>(ul, ur, wl, wr, b0, eos, build));
      // rotate: row j -> j-1, row j+1 -> j, Wl(j+1) pending, Wl(j)'s slot free
      const int nP = sQ, nQ = sF, nW = sP, nF = sW;
      sP = nP; sQ = nQ; sW = nW; sF = nF;
    }
  }
}

// ---------------------------------------------------------------------------
// integrate_emf3_corner (__integrate_2d__.c:632-702) at corner (j, i): the
// x1-flux of B2 is -E3 and the x2-flux of B1 is +E3; each of the four edge
// estimates is upwinded by the sign of the face mass flux (>0, <0, ==0).
// F1, F2: Cons1D flux groups (d = 0, By = 5, Bz = 6); ecc: cell-centred emf3.
__device__ __forceinline__ double corner_emf(const double* F1, const double* F2,
                                             const double* ecc, size_t ps, int pitch, int j,
                                             int i) {
  const uint32_t k = (uint32_t)(j * pitch + i);
  const uint32_t kjm = k - pitch, kim = k - 1, kjmim = k - pitch - 1;
  const double f1d_jm = LD(F1, kjm), f1d = LD(F1, k);
  const double f1by_jm = LD(F1 + 5 * ps, kjm), f1by = LD(F1 + 5 * ps, k);
  const double f2d_im = LD(F2, kim), f2d = LD(F2, k);
  const double f2bz_im = LD(F2 + 6 * ps, kim), f2bz = LD(F2 + 6 * ps, k);
  const double e_jmim = LD(ecc, kjmim), e_jm = LD(ecc, kjm), e_im = LD(ecc, kim), e_00 = LD(ecc, k);
  double emf_l1, emf_r1, emf_l2, emf_r2;
  if (f1d_jm > 0.0) emf_l2 = -f1by_jm + (f2bz_im - e_jmim);
  else if (f1d_jm < 0.0) emf_l2 = -f1by_jm + (f2bz - e_jm);
  else emf_l2 = -f1by_jm + 0.5 * (f2bz_im - e_jmim + f2bz - e_jm);
  if (f1d > 0.0) emf_r2 = -f1by + (f2bz_im - e_im);
  else if (f1d < 0.0) emf_r2 = -f1by + (f2bz - e_00);
  else emf_r2 = -f1by + 0.5 * (f2bz_im - e_im + f2bz - e_00);
  if (f2d_im > 0.0) emf_l1 = f2bz_im + (-f1by_jm - e_jmim);
  else if (f2d_im < 0.0) emf_l1 = f2bz_im + (-f1by - e_im);
  else emf_l1 = f2bz_im + 0.5 * (-f1by_jm - e_jmim - f1by - e_im);
  if (f2d > 0.0) emf_r1 = f2bz + (-f1by_jm - e_jm);
  else if (f2d < 0.0) emf_r1 = f2bz + (-f1by - e_00);
  else emf_r1 = f2bz + 0.5 * (-f1by_jm - e_jm - f1by - e_00);
  return 0.25 * (emf_l1 + emf_r1 + emf_l2 + emf_r2);
}

// ---------------------------------------------------------------------------
// integrate_emf3_corner (__integrate_2d__.c:632-702): upwinded corner emf from
// the face fluxes (x1-flux of B2 is -E3, x2-flux of B1 is +E3) and the
// cell-centred emf3. PRED: Step 4 over [js-1,je+2]x[is-1,ie+2] from predictor
// fluxes; otherwise Step 9 over [js,je+1]x[is,ie+1], the corners Step 9 uses.
template <bool PRED>
__global__ void __launch_bounds__(kTileX* kTileY) k_corner(const Inst* __restrict__ insts,
                                                            int ninst, int ntiles, int remap) {
  const int kind = PRED ? K_CORNER_P : K_CORNER_C;
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, kind, t)];
    if (!g.clk->active) continue;
    const int local = t - g.tile0[kind];
    const int w = PRED ? g.nx1 + 4 : g.nx1 + 2;
    const int ntx = dcdiv(w, kTileX);
    const int ty = local / ntx, tx = local - ty * ntx;
    const int i = (PRED ? g.is - 1 : g.is) + tx * kTileX + threadIdx.x;
    const int j = (PRED ? g.js - 1 : g.js) + ty * kTileY + threadIdx.y;
    if (i > g.ie + (PRED ? 2 : 1) || j > g.je + (PRED ? 2 : 1)) continue;
    const size_t ps = g.ps;
    const int pitch = g.pitch;
    double* __restrict__ sc = g.scratch;
    const double* F1 = sc + (PRED ? S_F1 : S_F1C) * ps;
    const double* F2 = sc + (PRED ? S_F2 : S_F2C) * ps;
    const double* ecc = sc + (PRED ? S_ECC : S_ECH) * ps;
    ST(sc + S_EMF * ps, (uint32_t)(j * pitch + i), corner_emf(F1, F2, ecc, ps, pitch, j, i));
  }
}

// ---------------------------------------------------------------------------
// Steps 4-8 for one cell (j,i) of [js-1,je+1]x[is-1,ie+1]:
//  - half-step face fields from the predictor corner emf (:266-275)
//  - x1-face states: transverse x2-flux gradients + MHD sources (:284-334),
//    then the corrector x1 flux (:438-447) for i >= is
//  - x2-face states: transverse x1-flux gradients + MHD sources (:343-393),
//    then the corrector x2 flux (:455-464) for j >= js
//  - d, M1, M2 at t^{n+1/2} and the cell-centred emf3 there (:402-431)
// PART 1: the x1 faces (Steps 5, 8b); PART 2: the x2 faces and Step 7 (6, 8c, 7).
// Split so that each thread evaluates one Riemann flux: the loads of the second
// face no longer compete for registers with the first face's flux.
// FLUX: the Riemann solver; a launch takes only the tiles of instances that use
// it. With one solver per kernel (the Roe one keeps its HLLE fallback) the
// corrector fits two waves per SIMD.
// CP (x1 part only): the tile first computes the Step 4 corner emf of its rows
// j0..j0+TY and columns i0..i0+kTileX into LDS from the predictor fluxes, and
// stores the corners it owns -- together the range [js-1,je+2]x[is-1,ie+2] --
// in the corner plane for the x2 part (no k_corner<pred> launch).
constexpr int kCornerCols = kTileX + 1;
template <int PART, int TY, int FLUX, bool CP = false>  // TY: rows of a tile of kTileX x TY faces
__global__ void __launch_bounds__(kTileX* TY, 2) k_corr(const Inst* __restrict__ insts,
                                                            int ninst, int ntiles, int remap,
                                                            int strip) {
  static_assert(!CP || PART == 1, "the corner phase belongs to the x1 part");
  __shared__ double park_lds[14 * kTileX * TY];
  __shared__ double cemf[CP ? (TY + 1) * kCornerCols : 1];
  const LdsPark park{park_lds, kTileX * TY, (int)(threadIdx.y * kTileX + threadIdx.x)};
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_CORR, t)];
    const Clock* clk = g.clk;
    if (!clk->active || g.flux != FLUX) continue;
    const int local = uni(t - g.tile0[K_CORR]);
    const int gis = uni(g.is), gjs = uni(g.js);
    int tx, ty;
    tile_pos(local, dcdiv(uni(g.nx1) + 3, kTileX), dcdiv(uni(g.nx2) + 3, TY), strip, tx, ty);
    const int i0 = gis - 1 + tx * kTileX, j0 = gjs - 1 + ty * TY;
    const int i = i0 + threadIdx.x;
    const int j = j0 + threadIdx.y;
    if (CP) {
      // every thread of the block reaches both barriers: the one before the
      // fill keeps the previous tile's readers ahead of it
      __syncthreads();
      double* __restrict__ scw = uni(g.scratch);
      const size_t psw = (size_t)uni((int)g.ps);
      const int pw = uni(g.pitch);
      const int je2 = uni(g.je) + 2, ie2 = uni(g.ie) + 2;
      for (int c = threadIdx.y * kTileX + threadIdx.x; c < (TY + 1) * kCornerCols;
           c += kTileX * TY) {
        const int cr = c / kCornerCols, cl = c - cr * kCornerCols;
        const int cj = j0 + cr, ci = i0 + cl;
        double e = 0.0;
        if (cj <= je2 && ci <= ie2) {
          e = corner_emf(scw + S_F1 * psw, scw + S_F2 * psw, scw + S_ECC * psw, psw, pw, cj, ci);
          // owned: the tile's own rows and columns, and the last row and
          // column of the range where they fall just outside them
          if ((cr < TY || cj == je2) && (cl < kTileX || ci == ie2))
            ST(scw + S_EMF * psw, (uint32_t)(cj * pw + ci), e);
        }
        cemf[c] = e;
      }
      __syncthreads();
    }
    if (i > uni(g.ie) + 1 || j > uni(g.je) + 1) continue;
    const double dt = uni(clk->dt);
    const double dtodx1 = dt / uni(g.dx1), dtodx2 = dt / uni(g.dx2);
    const double hdtodx1 = 0.5 * dtodx1, hdtodx2 = 0.5 * dtodx2;
    const EosC eos{uni(g.gamma), uni(g.g1), uni(g.g2)};
    const double* __restrict__ st = uni(g.state[clk->cur]);
    double* __restrict__ sc = uni(g.scratch);
    const size_t ps = (size_t)uni((int)g.ps);
    const int pitch = uni(g.pitch);
    const uint32_t k = (uint32_t)(j * pitch + i);
    const double* emf = sc + S_EMF * ps;
    const double* B1i = st + F_B1I * ps;
    const double* B2i = st + F_B2I * ps;
    auto b1h = [&](uint32_t kk) { return LD(B1i, kk) - hdtodx2 * (LD(emf, kk + pitch) - LD(emf, kk)); };
    auto b2h = [&](uint32_t kk) { return LD(B2i, kk) + hdtodx1 * (LD(emf, kk + 1) - LD(emf, kk)); };
    const double* F1 = sc + S_F1 * ps;
    const double* F2 = sc + S_F2 * ps;
    if (PART == 1 && i >= gis) {
      // corrected x1-face states (Steps 5a, 5b) at the half-step face field;
      // rebuilt from memory for the HLLE fallback
      auto build = [&](Cons7& L, Cons7& R, Prim7& wl, Prim7& wr, double& bx) {
        compiler_fence();
        L = get7(sc + S_UL1 * ps, ps, k);
        R = get7(sc + S_UR1 * ps, ps, k);
        {  // Step 5a: x2 fluxes at (j+1,i-1),(j,i-1) for L and (j+1,i),(j,i) for R
          const uint32_t a1 = k + pitch - 1, a0 = k - 1, c1 = k + pitch, c0 = k;
          L.d = L.d - hdtodx2 * (LD(F2, a1) - LD(F2, a0));
          L.mx = L.mx - hdtodx2 * (LD(F2 + 3 * ps, a1) - LD(F2 + 3 * ps, a0));
          L.my = L.my - hdtodx2 * (LD(F2 + 1 * ps, a1) - LD(F2 + 1 * ps, a0));
          L.mz = L.mz - hdtodx2 * (LD(F2 + 2 * ps, a1) - LD(F2 + 2 * ps, a0));
          L.e = L.e - hdtodx2 * (LD(F2 + 4 * ps, a1) - LD(F2 + 4 * ps, a0));
          L.bz = L.bz - hdtodx2 * (LD(F2 + 5 * ps, a1) - LD(F2 + 5 * ps, a0));
          R.d = R.d - hdtodx2 * (LD(F2, c1) - LD(F2, c0));
          R.mx = R.mx - hdtodx2 * (LD(F2 + 3 * ps, c1) - LD(F2 + 3 * ps, c0));
          R.my = R.my - hdtodx2 * (LD(F2 + 1 * ps, c1) - LD(F2 + 1 * ps, c0));
          R.mz = R.mz - hdtodx2 * (LD(F2 + 2 * ps, c1) - LD(F2 + 2 * ps, c0));
          R.e = R.e - hdtodx2 * (LD(F2 + 4 * ps, c1) - LD(F2 + 4 * ps, c0));
          R.bz = R.bz - hdtodx2 * (LD(F2 + 5 * ps, c1) - LD(F2 + 5 * ps, c0));
        }
        {  // Step 5b: MHD source terms with the cell-centred field of cells i-1, i
          double dbx = LD(B1i, k) - LD(B1i, k - 1);
          double b1 = LD(st + F_B1C * ps, k - 1), b2 = LD(st + F_B2C * ps, k - 1);
          double b3 = LD(st + F_B3C * ps, k - 1);
          double v3 = LD(st + F_M3 * ps, k - 1) / LD(st + F_D * ps, k - 1);
          L.mx = L.mx + hdtodx1 * b1 * dbx;
          L.my = L.my + hdtodx1 * b2 * dbx;
          L.mz = L.mz + hdtodx1 * b3 * dbx;
          L.bz = L.bz + hdtodx1 * v3 * dbx;
          L.e = L.e + hdtodx1 * b3 * v3 * dbx;
          dbx = LD(B1i, k + 1) - LD(B1i, k);
          b1 = LD(st + F_B1C * ps, k); b2 = LD(st + F_B2C * ps, k); b3 = LD(st + F_B3C * ps, k);
          v3 = LD(st + F_M3 * ps, k) / LD(st + F_D * ps, k);
          R.mx = R.mx + hdtodx1 * b1 * dbx;
          R.my = R.my + hdtodx1 * b2 * dbx;
          R.mz = R.mz + hdtodx1 * b3 * dbx;
          R.bz = R.bz + hdtodx1 * v3 * dbx;
          R.e = R.e + hdtodx1 * b3 * v3 * dbx;
        }
        if (CP) {
          const int cc = threadIdx.y * kCornerCols + threadIdx.x;
          bx = LD(B1i, k) - hdtodx2 * (cemf[cc + kCornerCols] - cemf[cc]);
        } else {
          bx = b1h(k);
        }
        wl = cons_to_prim(L, bx, eos);
        wr = cons_to_prim(R, bx, eos);
      };
      Cons7 L, R;
      Prim7 wl, wr;
      double bx;
      build(L, R, wl, wr, bx);
      put7(sc + S_F1C * ps, ps, k, riemann_of<FLUX>(L, R, wl, wr, bx, eos, build, park));
    }
    if (PART == 1) continue;
    if (j >= gjs) {
      // corrected x2-face states (Steps 6a, 6b) at the half-step face field
      auto build = [&](Cons7& L, Cons7& R, Prim7& wl, Prim7& wr, double& bx) {
        compiler_fence();
        L = get7(sc + S_UL2 * ps, ps, k);
        R = get7(sc + S_UR2 * ps, ps, k);
        {  // Step 6a: x1 fluxes at (j-1,i+1),(j-1,i) for L and (j,i+1),(j,i) for R
          const uint32_t a1 = k - pitch + 1, a0 = k - pitch, c1 = k + 1, c0 = k;
          L.d = L.d - hdtodx1 * (LD(F1, a1) - LD(F1, a0));
          L.mx = L.mx - hdtodx1 * (LD(F1 + 2 * ps, a1) - LD(F1 + 2 * ps, a0));
          L.my = L.my - hdtodx1 * (LD(F1 + 3 * ps, a1) - LD(F1 + 3 * ps, a0));
          L.mz = L.mz - hdtodx1 * (LD(F1 + 1 * ps, a1) - LD(F1 + 1 * ps, a0));
          L.e = L.e - hdtodx1 * (LD(F1 + 4 * ps, a1) - LD(F1 + 4 * ps, a0));
          L.by = L.by - hdtodx1 * (LD(F1 + 6 * ps, a1) - LD(F1 + 6 * ps, a0));
          R.d = R.d - hdtodx1 * (LD(F1, c1) - LD(F1, c0));
          R.mx = R.mx - hdtodx1 * (LD(F1 + 2 * ps, c1) - LD(F1 + 2 * ps, c0));
          R.my = R.my - hdtodx1 * (LD(F1 + 3 * ps, c1) - LD(F1 + 3 * ps, c0));
          R.mz = R.mz - hdtodx1 * (LD(F1 + 1 * ps, c1) - LD(F1 + 1 * ps, c0));
          R.e = R.e - hdtodx1 * (LD(F1 + 4 * ps, c1) - LD(F1 + 4 * ps, c0));
          R.by = R.by - hdtodx1 * (LD(F1 + 6 * ps, c1) - LD(F1 + 6 * ps, c0));
        }
        {  // Step 6b: MHD source terms with the cell-centred field of cells j-1, j
          const uint32_t km = k - pitch;
          double dby = LD(B2i, k) - LD(B2i, km);
          double b1 = LD(st + F_B1C * ps, km), b2 = LD(st + F_B2C * ps, km), b3 = LD(st + F_B3C * ps, km);
          double v3 = LD(st + F_M3 * ps, km) / LD(st + F_D * ps, km);
          L.mz = L.mz + hdtodx2 * b1 * dby;
          L.mx = L.mx + hdtodx2 * b2 * dby;
          L.my = L.my + hdtodx2 * b3 * dby;
          L.by = L.by + hdtodx2 * v3 * dby;
          L.e = L.e + hdtodx2 * b3 * v3 * dby;
          dby = LD(B2i, k + pitch) - LD(B2i, k);
          b1 = LD(st + F_B1C * ps, k); b2 = LD(st + F_B2C * ps, k); b3 = LD(st + F_B3C * ps, k);
          v3 = LD(st + F_M3 * ps, k) / LD(st + F_D * ps, k);
          R.mz = R.mz + hdtodx2 * b1 * dby;
          R.mx = R.mx + hdtodx2 * b2 * dby;
          R.my = R.my + hdtodx2 * b3 * dby;
          R.by = R.by + hdtodx2 * v3 * dby;
          R.e = R.e + hdtodx2 * b3 * v3 * dby;
        }
        bx = b2h(k);
        wl = cons_to_prim(L, bx, eos);
        wr = cons_to_prim(R, bx, eos);
      };
      Cons7 L, R;
      Prim7 wl, wr;
      double bx;
      build(L, R, wl, wr, bx);
      put7(sc + S_F2C * ps, ps, k, riemann_of<FLUX>(L, R, wl, wr, bx, eos, build, park));
    }
    // Step 7
    const double dhalf = LD(st + F_D * ps, k) - hdtodx1 * (LD(F1, k + 1) - LD(F1, k)) -
                         hdtodx2 * (LD(F2, k + pitch) - LD(F2, k));
    const double m1 = LD(st + F_M1 * ps, k) - hdtodx1 * (LD(F1 + 1 * ps, k + 1) - LD(F1 + 1 * ps, k)) -
                      hdtodx2 * (LD(F2 + 3 * ps, k + pitch) - LD(F2 + 3 * ps, k));
    const double m2 = LD(st + F_M2 * ps, k) - hdtodx1 * (LD(F1 + 2 * ps, k + 1) - LD(F1 + 2 * ps, k)) -
                      hdtodx2 * (LD(F2 + 1 * ps, k + pitch) - LD(F2 + 1 * ps, k));
    const double b1c = 0.5 * (b1h(k) + b1h(k + 1));
    const double b2c = 0.5 * (b2h(k) + b2h(k + pitch));
    ST(sc + S_ECH * ps, k, (b1c * m2 - b2c * m1) / dhalf);
  }
}

// ---------------------------------------------------------------------------
// new_dt.c:35-90 for one cell (the sound-speed form the reference keeps)
__device__ __forceinline__ double inverse_dt(const Inst& g, double d, double M1, double M2,
                                             double M3, double E) {
  const double di = 1.0 / d;
  const double v1 = M1 * di, v2 = M2 * di, v3 = M3 * di;
  const double qsq = v1 * v1 + v2 * v2 + v3 * v3;
  const double p = dmax(g.g1 * (E - 0.5 * d * qsq), kTiny);
  const double asq = g.gamma * p * di;
  double m = 0.0;
  if (g.nx1 > 1) m = dmax(m, (fabs(v1) + sqrt(asq)) / g.dx1);
  if (g.nx2 > 1) m = dmax(m, (fabs(v2) + sqrt(asq)) / g.dx2);
  return m;
}

// Health of one updated cell (FailBits): non-finite values, non-positive
// density, or a pressure below the floor that cons->prim would apply
// (convert_var.c:46-49 form: P = Gamma_1 (E - M^2/2d - B^2/2)).
__device__ __forceinline__ int cell_fail_bits(const Inst& g, double d, double M1, double M2,
                                              double M3, double E, double b1c, double b2c,
                                              double b3c) {
  int bits = 0;
  if (!(isfinite(d) && isfinite(M1) && isfinite(M2) && isfinite(M3) && isfinite(E) &&
        isfinite(b1c) && isfinite(b2c) && isfinite(b3c)))
    bits |= FAIL_NONFINITE;
  if (d <= 0.0) bits |= FAIL_DENSITY;
  const double p = g.g1 * (E - 0.5 * (M1 * M1 + M2 * M2 + M3 * M3) / d -
                           0.5 * (b1c * b1c + b2c * b2c + b3c * b3c));
  if (p < kTiny) bits |= FAIL_PRESSURE;
  return bits;
}

// Steps 9, 11a, 11b, 13 (:473-534) over [js,je+1]x[is,ie+1] into the other
// ping-pong buffer, and the new_dt maximum of the updated interior. The B2c
// update of 11a and the B1c update of 11b are dropped: Step 13 overwrites them.
// FUSED: the tile first computes the Step 9 corner emf it reads -- (TY+1)
// x (kTileX+1) corners -- into LDS from the corrector fluxes, instead of
// reading the plane k_corner<false> wrote (one launch and the corner plane's
// write and read fewer). TY: rows of a tile of kTileX x TY cells; the taller,
// the smaller the share of halo rows (corners and x2 fluxes of the row below)
// that the tile reads a second time.
template <bool FUSED, int TY>
__global__ void __launch_bounds__(kTileX* TY) k_final(const Inst* __restrict__ insts,
                                                       int ninst, int ntiles, int remap,
                                                       int strip) {
  __shared__ double red[TY];  // kTileX x TY threads: TY waves
  __shared__ double cemf[FUSED ? (TY + 1) * kCornerCols : 1];
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_FINAL, t)];
    Clock* clk = g.clk;
    if (!clk->active) continue;
    const int local = t - g.tile0[K_FINAL];
    int tx, ty;
    tile_pos(local, dcdiv(g.nx1 + 1, kTileX), dcdiv(g.nx2 + 1, TY), strip, tx, ty);
    const int i0 = g.is + tx * kTileX, j0 = g.js + ty * TY;
    const int i = i0 + threadIdx.x;
    const int j = j0 + threadIdx.y;
    if (FUSED) {
      // corners [js,je+1] x [is,ie+1] of this tile's rows j0..j0+TY and
      // columns i0..i0+kTileX (outside that range: never read)
      const double* sc = g.scratch;
      const size_t ps = g.ps;
      const int tid = threadIdx.y * kTileX + threadIdx.x;
      for (int c = tid; c < (TY + 1) * kCornerCols; c += kTileX * TY) {
        const int cj = j0 + c / kCornerCols, ci = i0 + c % kCornerCols;
        double e = 0.0;
        if (cj <= g.je + 1 && ci <= g.ie + 1)
          e = corner_emf(sc + S_F1C * ps, sc + S_F2C * ps, sc + S_ECH * ps, ps, g.pitch, cj, ci);
        cemf[c] = e;
      }
      __syncthreads();
    }
    double dti = 0.0;
    if (i <= g.ie + 1 && j <= g.je + 1) {
      const double dt = clk->dt;
      const double dtodx1 = dt / g.dx1, dtodx2 = dt / g.dx2;
      const int cur = clk->cur;
      const double* __restrict__ st = g.state[cur];
      double* __restrict__ out = g.state[cur ^ 1];
      const double* __restrict__ sc = g.scratch;
      const size_t ps = g.ps;
      const int pitch = g.pitch;
      const uint32_t k = (uint32_t)(j * pitch + i);
      const double* emf = sc + S_EMF * ps;
      const int cc = threadIdx.y * kCornerCols + threadIdx.x;
      auto corner = [&](int dj, int di) {
        return FUSED ? cemf[cc + dj * kCornerCols + di] : LD(emf, k + dj * pitch + di);
      };
      const double e00 = corner(0, 0), e10 = corner(1, 0), e01 = corner(0, 1);
      double b1n = 0.0, b2n = 0.0;
      if (j <= g.je) {
        b1n = LD(st + F_B1I * ps, k) - dtodx2 * (e10 - e00);
        ST(out + F_B1I * ps, k, b1n);
      }
      if (i <= g.ie) {
        b2n = LD(st + F_B2I * ps, k) + dtodx1 * (e01 - e00);
        ST(out + F_B2I * ps, k, b2n);
      }
      if (j <= g.je && i <= g.ie) {
        const double e11 = corner(1, 1);
        const double b1p = LD(st + F_B1I * ps, k + 1) - dtodx2 * (e11 - e01);
        const double b2p = LD(st + F_B2I * ps, k + pitch) + dtodx1 * (e11 - e10);
        const double* F1 = sc + S_F1C * ps;
        const double* F2 = sc + S_F2C * ps;
        double d = LD(st + F_D * ps, k), M1 = LD(st + F_M1 * ps, k), M2 = LD(st + F_M2 * ps, k);
        double M3 = LD(st + F_M3 * ps, k), E = LD(st + F_E * ps, k), B3c = LD(st + F_B3C * ps, k);
        d = d - dtodx1 * (LD(F1, k + 1) - LD(F1, k));
        M1 = M1 - dtodx1 * (LD(F1 + 1 * ps, k + 1) - LD(F1 + 1 * ps, k));
        M2 = M2 - dtodx1 * (LD(F1 + 2 * ps, k + 1) - LD(F1 + 2 * ps, k));
        M3 = M3 - dtodx1 * (LD(F1 + 3 * ps, k + 1) - LD(F1 + 3 * ps, k));
        E = E - dtodx1 * (LD(F1 + 4 * ps, k + 1) - LD(F1 + 4 * ps, k));
        B3c = B3c - dtodx1 * (LD(F1 + 6 * ps, k + 1) - LD(F1 + 6 * ps, k));
        const uint32_t kp = k + pitch;
        d = d - dtodx2 * (LD(F2, kp) - LD(F2, k));
        M1 = M1 - dtodx2 * (LD(F2 + 3 * ps, kp) - LD(F2 + 3 * ps, k));
        M2 = M2 - dtodx2 * (LD(F2 + 1 * ps, kp) - LD(F2 + 1 * ps, k));
        M3 = M3 - dtodx2 * (LD(F2 + 2 * ps, kp) - LD(F2 + 2 * ps, k));
        E = E - dtodx2 * (LD(F2 + 4 * ps, kp) - LD(F2 + 4 * ps, k));
        B3c = B3c - dtodx2 * (LD(F2 + 5 * ps, kp) - LD(F2 + 5 * ps, k));
        const double B1c = 0.5 * (b1n + b1p), B2c = 0.5 * (b2n + b2p);
        ST(out + F_D * ps, k, d);
        ST(out + F_M1 * ps, k, M1);
        ST(out + F_M2 * ps, k, M2);
        ST(out + F_M3 * ps, k, M3);
        ST(out + F_E * ps, k, E);
        ST(out + F_B1C * ps, k, B1c);
        ST(out + F_B2C * ps, k, B2c);
        ST(out + F_B3C * ps, k, B3c);
        ST(out + F_B3I * ps, k, B3c);
        dti = inverse_dt(g, d, M1, M2, M3, E);
        const int bits = cell_fail_bits(g, d, M1, M2, M3, E, B1c, B2c, B3c);
        if (bits) atomicOr(&clk->fail, bits);
      }
    }
    const double m = block_max(dti, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_pos(&clk->dti_slot[t % kDtSlots], m);
  }
}

// which = 0: fill the current state buffer; 1: fill the other one (the state
// being produced by this step); 2: carry the current buffer's ghost values
// unchanged into the other one (an in-place integrate_2d leaves them as they
// were). respect_active = 0 runs regardless of the clock.
__global__ void __launch_bounds__(kGhostBlock) k_ghost(const Inst* __restrict__ insts, int ninst,
                                                       int ntiles, int which, int respect_active,
                                                       int remap) {
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_GHOST, t)];
    const Clock* clk = g.clk;
    if (respect_active && !clk->active) continue;
    const int e = (t - g.tile0[K_GHOST]) * kGhostBlock + threadIdx.x;
    const int nrow = 8 * g.sx;
    if (e >= nrow + (g.nx2 + 1) * 8) continue;
    int j, i;
    if (e < nrow) {
      const int r = e / g.sx;
      i = e - r * g.sx;
      j = (r < 4) ? r : g.je + 1 + (r - 4);
    } else {
      const int e2 = e - nrow;
      const int r = e2 >> 3, c = e2 & 7;
      j = g.js + r;
      i = (c < 4) ? c : g.ie + 1 + (c - 4);
    }
    double* buf = g.state[which ? (clk->cur ^ 1) : clk->cur];
    const double* src = (which == 2) ? g.state[clk->cur] : buf;
    // where a stale source (a ghost's value before this fill, grids thinner
    // than the ghost layers) is read: filling the other buffer, the current
    // one holds the last fill's ghosts; filling in place, the launcher has
    // carried them into the other buffer first
    const double* pre = g.state[which == 1 ? clk->cur : (clk->cur ^ 1)];
    const size_t ps = g.ps;
    const uint32_t k = (uint32_t)(j * g.pitch + i);
    // the map once per plane class (cell-centred, B1i, B2i), not per plane
    int sj[3], si[3];
    bool gh[3], nx[3], ny[3], stale[3];
#pragma unroll
    for (int c = 0; c < 3; c++)
      gh[c] = ghost_chase(g, c, j, i, sj[c], si[c], nx[c], ny[c], stale[c]);
    // every plane's value is loaded before any is stored: a fill reads only
    // non-ghost cells of the buffer it writes (the map resolves chains to
    // them) or the other buffer, so no store can alias a later load, and the
    // eleven loads are in flight together
    double v[NSTATE];
#pragma unroll
    for (int f = 0; f < NSTATE; f++) {
      const int c = ghost_class(f);
      v[f] = 0.0;
      if (gh[c]) {
        if (which == 2) {
          v[f] = LD(src + f * ps, k);
        } else {
          const double x = LD((stale[c] ? pre : src) + f * ps, (size_t)sj[c] * g.pitch + si[c]);
          v[f] = ghost_negates(f, nx[c], ny[c]) ? -x : x;
        }
      }
    }
#pragma unroll
    for (int f = 0; f < NSTATE; f++)
      if (gh[ghost_class(f)]) ST(buf + f * ps, k, v[f]);
  }
}

// ---------------------------------------------------------------------------
// Clock kernels: one workgroup of 256 threads per instance.
// mode 0: main-loop step (main.c:221-232): run while time < tlim and
//         (nlim < 0 or nstep < nlim) and no failure bit of fail_stop is set;
//         clamp dt to reach tlim exactly. The
//         step also waits (does not run) once time >= tpause, so that the
//         host can write the outputs due at the top of the iteration
//         (main.c:226, output.c:372-376) before it runs.
// mode 1: unconditional single integration (drop-in integrate_2d).
__global__ void __launch_bounds__(256) k_begin(const Inst* __restrict__ insts, int mode) {
  Clock* c = insts[blockIdx.x].clk;
  for (int s = threadIdx.x; s < kDtSlots; s += blockDim.x) c->dti_slot[s] = 0ull;
  if (threadIdx.x == 0) {
    int act = 1;
    if (mode == 0) {
      act = (c->time < c->tlim) && (c->nlim < 0 || c->nstep < c->nlim) &&
            (c->time < c->tpause) && ((c->fail & c->fail_stop) == 0);
      if (act && (c->tlim - c->time) < c->dt) c->dt = (c->tlim - c->time);
    }
    c->active = act;
  }
}

__device__ __forceinline__ double reduce_slots(const Clock* c, double* red) {
  double m = __longlong_as_double((long long)c->dti_slot[threadIdx.x]);
  return block_max(m, red);
}

// mode 0 (end of a main-loop step, main.c:289-292 + new_dt.c:93-97):
//   nstep++, time += dt, dt = MIN(2 dt, CourNo/max_dti), swap buffers
// mode 1 (end of an unconditional integration): swap buffers only
// Both record the step number of the first state that raised a failure bit.
__global__ void __launch_bounds__(kDtSlots) k_end(const Inst* __restrict__ insts, int mode) {
  __shared__ double red[4];
  Clock* c = insts[blockIdx.x].clk;
  const double max_dti = reduce_slots(c, red);
  if (threadIdx.x == 0 && c->active) {
    if (mode == 0) {
      c->nstep = c->nstep + 1;
      c->time = c->time + c->dt;
      const double a = 2.0 * c->dt, b = c->cour_no / max_dti;
      c->dt = (c->nstep == 0) ? b : dmin(a, b);
      c->steps_run = c->steps_run + 1;
    }
    if (c->fail != 0 && c->fail_nstep < 0) c->fail_nstep = mode == 0 ? c->nstep : c->nstep + 1;
    c->cur = c->cur ^ 1;
  }
}

// standalone new_dt (new_dt.c:25-103) on the current state
__global__ void __launch_bounds__(kTileX* kTileY) k_dti(const Inst* __restrict__ insts, int ninst,
                                                         int ntiles, int remap) {
  __shared__ double red[4];
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_INTERIOR, t)];
    Clock* clk = g.clk;
    const int local = t - g.tile0[K_INTERIOR];
    const int ntx = dcdiv(g.nx1, kTileX);
    const int ty = local / ntx, tx = local - ty * ntx;
    const int i = g.is + tx * kTileX + threadIdx.x;
    const int j = g.js + ty * kTileY + threadIdx.y;
    double dti = 0.0;
    if (i <= g.ie && j <= g.je) {
      const double* st = g.state[clk->cur];
      const size_t ps = g.ps;
      const uint32_t k = (uint32_t)(j * g.pitch + i);
      dti = inverse_dt(g, LD(st + F_D * ps, k), LD(st + F_M1 * ps, k), LD(st + F_M2 * ps, k),
                       LD(st + F_M3 * ps, k), LD(st + F_E * ps, k));
    }
    const double m = block_max(dti, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_pos(&clk->dti_slot[t % kDtSlots], m);
  }
}

__global__ void __launch_bounds__(kDtSlots) k_dt_final(const Inst* __restrict__ insts) {
  __shared__ double red[4];
  Clock* c = insts[blockIdx.x].clk;
  const double max_dti = reduce_slots(c, red);
  if (threadIdx.x == 0) {
    const double b = c->cour_no / max_dti;
    c->dt = (c->nstep == 0) ? b : dmin(2.0 * c->dt, b);
  }
}

__global__ void __launch_bounds__(256) k_clear_slots(const Inst* __restrict__ insts) {
  Clock* c = insts[blockIdx.x].clk;
  for (int s = threadIdx.x; s < kDtSlots; s += blockDim.x) c->dti_slot[s] = 0ull;
  if (threadIdx.x == 0) c->divb_bits = 0ull;
}

// max over the interior of |div B| = |dB1/dx1 + dB2/dx2| on the current state
__global__ void __launch_bounds__(kTileX* kTileY) k_divb(const Inst* __restrict__ insts, int ninst,
                                                          int ntiles, int remap) {
  __shared__ double red[4];
  for (int t0 = blockIdx.x; t0 < ntiles; t0 += gridDim.x) {
    const int t = xcd_tile(t0, ntiles, remap);
    const Inst& g = insts[find_inst(insts, ninst, K_INTERIOR, t)];
    Clock* clk = g.clk;
    const int local = t - g.tile0[K_INTERIOR];
    const int ntx = dcdiv(g.nx1, kTileX);
    const int ty = local / ntx, tx = local - ty * ntx;
    const int i = g.is + tx * kTileX + threadIdx.x;
    const int j = g.js + ty * kTileY + threadIdx.y;
    double v = 0.0;
    if (i <= g.ie && j <= g.je) {
      const double* st = g.state[clk->cur];
      const size_t ps = g.ps;
      const uint32_t k = (uint32_t)(j * g.pitch + i);
      v = fabs((LD(st + F_B1I * ps, k + 1) - LD(st + F_B1I * ps, k)) / g.dx1 +
               (LD(st + F_B2I * ps, k + g.pitch) - LD(st + F_B2I * ps, k)) / g.dx2);
    }
    const double m = block_max(v, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_pos(&clk->divb_bits, m);
  }
}

// ---------------------------------------------------------------------------
// Host-side launchers of this variant
// resident x1-predictor tiles per CU at two waves per SIMD
constexpr int kX1PerCu = 2 * 4 * 64 / kX1Block;

typedef void (*PredX1Kernel)(const Inst*, int, int, int);
typedef void (*PredX2Kernel)(const Inst*, int, int, int, int);
static PredX1Kernel pred_x1_kernel(bool save, int flux) {
  if (flux == kFluxHlle) return save ? k_pred_x1<true, kFluxHlle> : k_pred_x1<false, kFluxHlle>;
  return save ? k_pred_x1<true, kFluxRoe> : k_pred_x1<false, kFluxRoe>;
}
static PredX2Kernel pred_x2_kernel(bool save, int flux, int kind) {
  if (kind == 1) {
    if (flux == kFluxHlle)
      return save ? k_pred_x2l<true, kFluxHlle> : k_pred_x2l<false, kFluxHlle>;
    return save ? k_pred_x2l<true, kFluxRoe> : k_pred_x2l<false, kFluxRoe>;
  }
  if (flux == kFluxHlle) return save ? k_pred_x2<true, kFluxHlle> : k_pred_x2<false, kFluxHlle>;
  return save ? k_pred_x2<true, kFluxRoe> : k_pred_x2<false, kFluxRoe>;
}
typedef void (*CorrKernel)(const Inst*, int, int, int, int);
typedef void (*FinalKernel)(const Inst*, int, int, int, int);
// the corrector instantiation for a part (1: x1 faces, 2: x2 faces and Step
// 7), tile height and Riemann solver
template <int PART, bool CP>
static CorrKernel corr_kernel_part(int ty, int flux) {
  if (flux == kFluxHlle)
    return ty == 1 ? k_corr<PART, 1, kFluxHlle, CP>
                   : ty == 2 ? k_corr<PART, 2, kFluxHlle, CP> : k_corr<PART, 4, kFluxHlle, CP>;
  return ty == 1 ? k_corr<PART, 1, kFluxRoe, CP>
                 : ty == 2 ? k_corr<PART, 2, kFluxRoe, CP> : k_corr<PART, 4, kFluxRoe, CP>;
}
// cp: the x1 part computes the Step 4 corners itself (fusion depth 2)
static CorrKernel corr_kernel(int part, int ty, int flux, bool cp) {
  if (part == 2) return corr_kernel_part<2, false>(ty, flux);
  return cp ? corr_kernel_part<1, true>(ty, flux) : corr_kernel_part<1, false>(ty, flux);
}

// persistent grids: persist x per_cu workgroups per CU (persist 0: one per tile)
static int grid_of(const BatchLaunch& b, int kind, int per_cu) {
  const int n = b.ntiles[kind];
  if (b.shape.persist <= 0) return n;
  const int cap = b.ncu * b.shape.persist * per_cu;
  return n < cap ? n : cap;
}

}  // namespace ATHMI_VARIANT
}  // namespace athmi

#define ATHMI_CAT2(a, b) a##b
#define ATHMI_CAT(a, b) ATHMI_CAT2(a, b)
#define ATHMI_FN(name) ATHMI_CAT(ATHMI_CAT(athmi_, ATHMI_VARIANT), _##name)

using namespace athmi;
using namespace athmi::ATHMI_VARIANT;

extern "C" {

// Launch one kernel of the pipeline for every instance of the batch.
hipError_t ATHMI_FN(launch)(const void* bl, int kernel, hipStream_t s) {
  const BatchLaunch& b = *static_cast<const BatchLaunch*>(bl);
  const dim3 b2(kTileX, kTileY);
  switch (kernel) {
    case SK_BEGIN_STEP:
    case SK_BEGIN_ONCE:
      hipLaunchKernelGGL(k_begin, dim3(b.ninst), dim3(256), 0, s, b.insts,
                         kernel == SK_BEGIN_STEP ? 0 : 1);
      break;
    case SK_PRED_X1:
    case SK_PRED_X1_SAVE:
      for (int f = 0; f < 2; f++) {
        if (!(b.flux_mask & (1 << f))) continue;
        const PredX1Kernel fn = pred_x1_kernel(kernel == SK_PRED_X1_SAVE, f);
        hipLaunchKernelGGL(fn, dim3(grid_of(b, K_PRED_X1, kX1PerCu)), dim3(kX1Block), 0, s,
                           b.insts, b.ninst, b.ntiles[K_PRED_X1], b.shape.xcd_remap);
      }
      break;
    case SK_PRED_X2:
    case SK_PRED_X2_SAVE:
      for (int f = 0; f < 2; f++) {
        if (!(b.flux_mask & (1 << f))) continue;
        const int kind = b.shape.x2_kind;
        const PredX2Kernel fn = pred_x2_kernel(kernel == SK_PRED_X2_SAVE, f, kind);
        hipLaunchKernelGGL(fn, dim3(grid_of(b, K_PRED_X2, kind == 1 ? 8 : 16)), dim3(kX2Block),
                           0, s, b.insts,
                           b.ninst, b.ntiles[K_PRED_X2], b.shape.x2_seg, b.shape.xcd_remap);
      }
      break;
    case SK_CORNER_P:
      // at fusion depth 2 the x1 corrector computes these corners
      if (b.shape.fuse_corner >= 2) break;
      hipLaunchKernelGGL(k_corner<true>, dim3(grid_of(b, K_CORNER_P, 8)), b2, 0, s, b.insts,
                         b.ninst, b.ntiles[K_CORNER_P], b.shape.xcd_remap);
      break;
    case SK_CORR:
    case SK_CORR2: {
      // the same threads in flight per CU whatever the tile height; one launch
      // per Riemann solver in use
      const int ty = b.shape.corr_rows;
      const dim3 grid(grid_of(b, K_CORR, 4 * (kTileY / ty))), blk(kTileX, ty);
      const bool x1 = (kernel == SK_CORR);
      for (int f = 0; f < 2; f++) {
        if (!(b.flux_mask & (1 << f))) continue;
        const CorrKernel fn = corr_kernel(x1 ? 1 : 2, ty, f, b.shape.fuse_corner >= 2);
        hipLaunchKernelGGL(fn, grid, blk, 0, s, b.insts, b.ninst, b.ntiles[K_CORR],
                           b.shape.xcd_remap, b.shape.strip);
      }
      break;
    }
    case SK_CORNER_C:
      // with the fused final kernel the corners are computed there
      if (b.shape.fuse_corner) break;
      hipLaunchKernelGGL(k_corner<false>, dim3(grid_of(b, K_CORNER_C, 8)), b2, 0, s, b.insts,
                         b.ninst, b.ntiles[K_CORNER_C], b.shape.xcd_remap);
      break;
    case SK_FINAL: {
      // 32 waves per CU whatever the tile height
      const int ty = final_rows_of(b.shape.final_rows);
      const bool f = b.shape.fuse_corner != 0;
      const FinalKernel fn = ty == 16 ? (f ? k_final<true, 16> : k_final<false, 16>)
                             : ty == 8 ? (f ? k_final<true, 8> : k_final<false, 8>)
                                       : (f ? k_final<true, 4> : k_final<false, 4>);
      hipLaunchKernelGGL(fn, dim3(grid_of(b, K_FINAL, 32 / ty)), dim3(kTileX, ty), 0, s, b.insts,
                         b.ninst, b.ntiles[K_FINAL], b.shape.xcd_remap, b.shape.strip);
      break;
    }
    case SK_GHOST_CUR:
    case SK_GHOST_NEXT:
    case SK_GHOST_CARRY: {
      const int which = kernel == SK_GHOST_CUR ? 0 : (kernel == SK_GHOST_NEXT ? 1 : 2);
      // an in-place fill that reads stale ghosts reads them from the other
      // buffer: carry them there first
      if (which == 0 && b.stale_ghosts)
        hipLaunchKernelGGL(k_ghost, dim3(grid_of(b, K_GHOST, 8)), dim3(kGhostBlock), 0, s,
                           b.insts, b.ninst, b.ntiles[K_GHOST], 2, 0, b.shape.xcd_remap);
      hipLaunchKernelGGL(k_ghost, dim3(grid_of(b, K_GHOST, 8)), dim3(kGhostBlock), 0, s, b.insts,
                         b.ninst, b.ntiles[K_GHOST], which, kernel == SK_GHOST_NEXT ? 1 : 0,
                         b.shape.xcd_remap);
      break;
    }
    case SK_END_STEP:
    case SK_END_ONCE:
      hipLaunchKernelGGL(k_end, dim3(b.ninst), dim3(kDtSlots), 0, s, b.insts,
                         kernel == SK_END_STEP ? 0 : 1);
      break;
    case SK_CLEAR:
      hipLaunchKernelGGL(k_clear_slots, dim3(b.ninst), dim3(256), 0, s, b.insts);
      break;
    case SK_DTI:
      hipLaunchKernelGGL(k_dti, dim3(grid_of(b, K_INTERIOR, 8)), b2, 0, s, b.insts, b.ninst,
                         b.ntiles[K_INTERIOR], b.shape.xcd_remap);
      break;
    case SK_DT_FINAL:
      hipLaunchKernelGGL(k_dt_final, dim3(b.ninst), dim3(kDtSlots), 0, s, b.insts);
      break;
    case SK_DIVB:
      hipLaunchKernelGGL(k_divb, dim3(grid_of(b, K_INTERIOR, 8)), b2, 0, s, b.insts, b.ninst,
                         b.ntiles[K_INTERIOR], b.shape.xcd_remap);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // extern "C"
