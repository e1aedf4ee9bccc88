// ghost_map.hpp -- the set_bvals ghost-cell map shared by the forward ghost
// fill and its adjoint (which folds each ghost cell's adjoint back into the
// cell it copies).
#pragma once
#include <hip/hip_runtime.h>

#include "step_kernels.hpp"

namespace athmi {

// ---------------------------------------------------------------------------
// set_bvals (set_bvals.c:114-145): the x1 fills then the x2 fills compose into
// one map from each ghost cell to the interior cell it copies, so a single
// pass reads only non-ghost cells. Per plane, the exceptions of the reference
// hold: LD(B1i, ie+1) and LD(B2i, je+1) are not boundary values (:630, :873, :1011);
// B2i's x1 fills cover rows js..je+1; reflection negates M1 across x1 and M2
// across x2 (:312-545), outflow copies the edge cell (:553-786), periodic
// wraps (:793-1031).
__device__ __forceinline__ bool ghost_src(const Inst& g, int f, int j, int i, int& sj, int& si,
                                          bool& neg) {
  const bool b1 = (f == F_B1I), b2 = (f == F_B2I);
  const int ie_x = b1 ? g.ie + 1 : g.ie;  // last non-ghost column of this plane
  const int je_y = b2 ? g.je + 1 : g.je;  // last non-ghost row
  const bool gx = (i < g.is) || (i > ie_x);
  const bool gy = (j < g.js) || (j > je_y);
  bool nx = false, ny = false;
  auto map_x = [&](int ii) -> int {
    if (ii < g.is) {
      const int kk = g.is - ii;
      if (g.bc[0] == BC_PERIODIC) return g.ie - (kk - 1);
      if (g.bc[0] == BC_OUTFLOW) return g.is;
      nx = true;
      return g.is + (kk - 1);
    }
    const int kk = ii - g.ie;
    if (g.bc[1] == BC_PERIODIC) return g.is + (kk - 1);
    if (g.bc[1] == BC_OUTFLOW) return g.ie;
    nx = true;
    return b1 ? g.ie - (kk - 2) : g.ie - (kk - 1);
  };
  auto map_y = [&](int jj) -> int {
    if (jj < g.js) {
      const int kk = g.js - jj;
      if (g.bc[2] == BC_PERIODIC) return g.je - (kk - 1);
      if (g.bc[2] == BC_OUTFLOW) return g.js;
      ny = true;
      return g.js + (kk - 1);
    }
    const int kk = jj - g.je;
    if (g.bc[3] == BC_PERIODIC) return g.js + (kk - 1);
    if (g.bc[3] == BC_OUTFLOW) return g.je;
    ny = true;
    return b2 ? g.je - (kk - 2) : g.je - (kk - 1);
  };
  if (gy) {
    sj = map_y(j);
    si = gx ? map_x(i) : i;
  } else if (gx && j >= g.js && j <= je_y) {
    sj = j;
    si = map_x(i);
  } else {
    return false;
  }
  neg = (f == F_M1 && nx) || (f == F_M2 && ny);
  return true;
}

}  // namespace athmi
