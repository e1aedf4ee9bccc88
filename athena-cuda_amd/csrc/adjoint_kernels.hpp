// adjoint_kernels.hpp -- layout shared by the host side and the adjoint
// (reverse-mode) kernels of one CTU+CT step of a single instance.
#pragma once
#include <hip/hip_runtime.h>

#include "step_kernels.hpp"

namespace athmi {

// adjoint planes (stride Inst::ps, same pitch as the state); each Cons1D-shaped
// group is 7 planes in the sweep's own order
enum AdjPlane {
  A_F1C = 0,    // corrector x1 fluxes
  A_F2C = 7,    // corrector x2 fluxes
  A_UL1 = 14,   // x1-face states (predictor value = corrected value's identity part)
  A_UR1 = 21,
  A_UL2 = 28,   // x2-face states
  A_UR2 = 35,
  A_F1 = 42,    // predictor x1 fluxes
  A_F2 = 49,    // predictor x2 fluxes
  A_WL1 = 56,   // PLM outputs (primitive) at x1 faces before the source terms
  A_WR1 = 63,
  A_WL2 = 70,   // ... at x2 faces
  A_WR2 = 77,
  A_B1H = 84,   // half-step face fields
  A_B2H = 85,
  A_EMFC = 86,  // corner emf, corrector
  A_EMFP = 87,  // corner emf, predictor
  A_ECH = 88,   // cell-centred emf3 at t^{n+1/2}
  A_ECC = 89,   // cell-centred emf3 at t^n
  NADJ = 90
};

// Everything the adjoint kernels of one step need
struct AdjArgs {
  Inst g;              // the instance (geometry, EOS, pitch, plane stride)
  const double* st;    // S^n: the state the forward pass of this step started from
  const double* sc;    // the forward scratch of this step (predictor + corrector)
  double* ab;          // NADJ adjoint planes
  double* ain;         // adjoint of S^n (NSTATE planes), accumulated
  double* aout;        // adjoint of S^{n+1} (NSTATE planes)
  double dt;
  int x2_seg;          // faces per lane of the x2-predictor adjoint
};

enum AdjKernel {
  AK_FOLD_OUT = 0,  // adjoint of set_bvals on aout (ghosts folded into their sources)
  AK_FOLD_IN,       // ... on ain
  AK_ZERO,          // clear ain and the adjoint planes
  AK_UPDATE,        // Steps 9, 11a, 11b, 13
  AK_CORNER_C,      // Step 9 corner emf
  AK_RIEMANN_C,     // Steps 8b, 8c (corrector fluxes)
  AK_HALF,          // Step 7 (d, M1, M2 and emf3 at t^{n+1/2})
  AK_TRANSVERSE,    // Steps 5, 6 and the half-step face fields of Step 4
  AK_CORNER_P,      // Step 4 corner emf
  AK_EMF_CC,        // Step 3
  AK_PRED_X1,       // Steps 1c-1e (sources, prim->cons, predictor flux) per x1 face
  AK_PLM_X1,        // Steps 1a-1b (cons->prim, PLM) per cell
  AK_PRED_X2,       // Steps 2c-2e per x2 face
  AK_PLM_X2,        // Steps 2a-2b per cell
  AK_COUNT
};

}  // namespace athmi

extern "C" hipError_t athmi_adjoint_launch(const athmi::AdjArgs* a, int kernel, hipStream_t s);
