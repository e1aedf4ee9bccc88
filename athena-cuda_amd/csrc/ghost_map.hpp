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
__device__ __forceinline__ bool ghost_src(const Inst& g, int f, int j, int i, int& sj, int& si,
                                          bool& neg, bool& stale) {
  const bool b1 = (f == F_B1I), b2 = (f == F_B2I);
  const int ie_x = b1 ? g.ie + 1 : g.ie;  // last non-ghost column of this plane
  const int je_y = b2 ? g.je + 1 : g.je;  // last non-ghost row
  const bool gx = (i < g.is) || (i > ie_x);
  const bool gy = (j < g.js) || (j > je_y);
  if (!gx && !gy) return false;
  bool nx = false, ny = false;
  stale = false;
  // the column whose value ghost column c of a non-ghost row ends with
  auto chase_x = [&](int c) -> int {
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
  neg = (f == F_M1 && nx) || (f == F_M2 && ny);
  return true;
}

// the map for grids of Nx1, Nx2 >= 4, where no source is stale
__device__ __forceinline__ bool ghost_src(const Inst& g, int f, int j, int i, int& sj, int& si,
                                          bool& neg) {
  bool stale;
  return ghost_src(g, f, j, i, sj, si, neg, stale);
}

}  // namespace athmi
