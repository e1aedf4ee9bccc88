// ghost_map.hpp -- the set_bvals ghost-cell map shared by the forward ghost
// fill and its adjoint (which folds each ghost cell's adjoint back into the
// cell it copies).
#pragma once
#include <hip/hip_runtime.h>

#include "step_kernels.hpp"

namespace athmi {

// ---------------------------------------------------------------------------
// set_bvals (set_bvals.c:114-145) fills in place and in order: the inner x1
// fill (ghost columns is-1, is-2, ... in that order), the outer x1 fill, then
// the inner and outer x2 fills over whole rows (set_bvals.c:312-1031). The
// fills compose here into one map from each ghost cell to the cell whose
// value it ends up with, so that one parallel pass reproduces them. Per plane
// the exceptions of the reference hold: B1i at ie+1 and B2i at je+1 are not
// boundary values (:630, :873, :1011); B2i's x1 fills cover rows js..je+1;
// reflection negates M1 across x1 and M2 across x2 (:312-545), outflow copies
// the edge cell (:553-786), periodic wraps (:793-1031).
// With Nx >= 4 every source is a non-ghost cell. A thinner grid copies ghost
// cells too: a periodic wrap or an outer reflecting fill copies a ghost the
// same set_bvals wrote earlier (the map follows it to where that value came
// from), and an inner reflecting fill copies an outer ghost the outer fill has
// not reached yet, i.e. the value it held before this set_bvals (stale).
// The map depends on the plane only through its class: cell-centred (0), B1i
// (1) or B2i (2). ghost_chase gives the source cell of ghost (j, i) of a class,
// whether the x1 (nx) and x2 (ny) fills reflected it an odd number of times,
// and whether the source is stale; false if (j, i) is not a ghost of the class.
__device__ __forceinline__ bool ghost_chase(const Inst& g, int cls, int j, int i, int& sj,
                                            int& si, bool& nx, bool& ny, bool& stale) {
  const bool b1 = (cls == 1), b2 = (cls == 2);
  const int ie_x = b1 ? g.ie + 1 : g.ie;  // last non-ghost column of this plane
  const int je_y = b2 ? g.je + 1 : g.je;  // last non-ghost row
  const bool gx = (i < g.is) || (i > ie_x);
  const bool gy = (j < g.js) || (j > je_y);
  nx = false;
  ny = false;
  stale = false;
  if (!gx && !gy) return false;
  // the column whose value ghost column c of a non-ghost row ends with
  // (the hops stay a loop: unrolled, the eleven planes' copies of the map
  // made the fill kernel larger than the instruction cache)
  auto chase_x = [&](int c) -> int {
#pragma unroll 1
    for (int hop = 0; hop < 8 && (c < g.is || c > ie_x); hop++) {
      if (c < g.is) {
        const int kk = g.is - c;
        if (g.bc[0] == BC_PERIODIC) {
          c = g.ie - (kk - 1);  // a ghost only if Nx1 < kk: written before, by this loop
        } else if (g.bc[0] == BC_OUTFLOW) {
          c = g.is;
        } else {
          nx = !nx;
          c = g.is + (kk - 1);
          if (c > ie_x) {  // an outer ghost, before the outer fill
            stale = true;
            return c;
          }
        }
      } else {
        const int kk = c - g.ie;
        if (g.bc[1] == BC_PERIODIC) {
          c = g.is + (kk - 1);  // a ghost only if written before, by this loop
        } else if (g.bc[1] == BC_OUTFLOW) {
          c = g.ie;
        } else {
          nx = !nx;
          c = b1 ? g.ie - (kk - 2) : g.ie - (kk - 1);  // an inner ghost is already filled
        }
      }
    }
    return c;
  };
  // the same along x2 for ghost row r (whole rows, after the x1 fills)
  auto chase_y = [&](int r) -> int {
#pragma unroll 1
    for (int hop = 0; hop < 8 && (r < g.js || r > je_y); hop++) {
      if (r < g.js) {
        const int kk = g.js - r;
        if (g.bc[2] == BC_PERIODIC) {
          r = g.je - (kk - 1);
        } else if (g.bc[2] == BC_OUTFLOW) {
          r = g.js;
        } else {
          ny = !ny;
          r = g.js + (kk - 1);
          if (r > je_y) {
            stale = true;
            return r;
          }
        }
      } else {
        const int kk = r - g.je;
        if (g.bc[3] == BC_PERIODIC) {
          r = g.js + (kk - 1);
        } else if (g.bc[3] == BC_OUTFLOW) {
          r = g.je;
        } else {
          ny = !ny;
          r = b2 ? g.je - (kk - 2) : g.je - (kk - 1);
        }
      }
    }
    return r;
  };
  if (gy) {
    sj = chase_y(j);
    // a stale row was not touched by the x1 fills either; otherwise the row
    // is a non-ghost row whose ghost columns hold the x1 fills' values
    si = (!stale && gx) ? chase_x(i) : i;
  } else {
    sj = j;
    si = chase_x(i);
  }
  return true;
}

__device__ __forceinline__ constexpr int ghost_class(int f) {
  return f == F_B1I ? 1 : (f == F_B2I ? 2 : 0);
}

// reflection negates M1 across x1 and M2 across x2
__device__ __forceinline__ constexpr bool ghost_negates(int f, bool nx, bool ny) {
  return (f == F_M1 && nx) || (f == F_M2 && ny);
}

// the map of plane f
__device__ __forceinline__ bool ghost_src(const Inst& g, int f, int j, int i, int& sj, int& si,
                                          bool& neg, bool& stale) {
  bool nx, ny;
  const bool ghost = ghost_chase(g, ghost_class(f), j, i, sj, si, nx, ny, stale);
  neg = ghost_negates(f, nx, ny);
  return ghost;
}

// the map for grids of Nx1, Nx2 >= 4, where no source is stale
__device__ __forceinline__ bool ghost_src(const Inst& g, int f, int j, int i, int& sj, int& si,
                                          bool& neg) {
  bool stale;
  return ghost_src(g, f, j, i, sj, si, neg, stale);
}

}  // namespace athmi
