// adjoint_kernels.hpp -- layout shared by the host side and the adjoint
// (reverse-mode) kernels of one CTU+CT step of a single instance.
#pragma once
#include <hip/hip_runtime.h>

#include "step_kernels.hpp"

namespace athmi {

// adjoint planes (stride Inst::ps, same pitch as the state); each Cons1D-shaped
// group is 7 planes in the sweep's own order. The first NACC planes are
// accumulated by several kernels and cleared every step; each of the others
// is stored by one kernel over the range its readers cover (entries outside
// that range stay zero from the clear at the start of a reverse pass).
enum AdjPlane {
  A_ECH = 0,    // cell-centred emf3 at t^{n+1/2}
  A_ECC = 1,    // cell-centred emf3 at t^n
  NACC = 2,
  A_F1 = 2,     // predictor x1 fluxes
  A_F2 = 9,     // predictor x2 fluxes
  A_B1H = 16,   // half-step face fields, from the corrector Riemann fluxes
  A_B2H = 17,
  A_EMFP = 18,  // corner emf, predictor
  A_F1C = 19,   // corrector x1 fluxes
  A_F2C = 26,   // corrector x2 fluxes
  A_UL1 = 33,   // x1-face states (predictor value = corrected value's identity part)
  A_UR1 = 40,
  A_UL2 = 47,   // x2-face states
  A_UR2 = 54,
  A_WL1 = 61,   // PLM outputs (primitive) at x1 faces before the source terms
  A_WR1 = 68,
  A_WL2 = 75,   // ... at x2 faces
  A_WR2 = 82,
  A_EMFC = 89,  // corner emf, corrector
  A_HD = 90,    // Step 7 per cell: half-step d, M1, M2
  A_HM1 = 91,
  A_HM2 = 92,
  A_HB1 = 93,   // ... and the cell-centred half-step B1c, B2c
  A_HB2 = 94,
  A_PW1 = 95,   // PLM reverse per cell (x1 sweep): adjoints of its 3 stencil cells'
                // primitive states (3 x 7) and of its Bxc (1)
  A_PW2 = 117,  // ... x2 sweep
  A_PC1 = 139,  // deterministic reverse: the predictor-face source adjoints of the x1
                // sweep, per face: B(k-1), B(k), B(k+1), M(k-1), d(k-1), M(k), d(k)
  A_PC2 = 146,  // ... x2 sweep
  NADJ = 153
};

// Everything the adjoint kernels of one step need
struct AdjArgs {
  Inst g;              // the instance (geometry, EOS, pitch, plane stride)
  const double* st;    // S^n: the state the forward pass of this step started from
  const double* sc;    // the forward scratch of this step (predictor + corrector)
  const double* emfp;  // the step's predictor corner emf (Step 4)
  const double* pw;    // NPRIMS planes: the predictor's primitive face states
  double* ab;          // NADJ adjoint planes
  double* ain;         // adjoint of S^n (NSTATE planes), accumulated
  double* aout;        // adjoint of S^{n+1} (NSTATE planes)
  double dt;
  int det;             // 1: deterministic reverse (every sum in a fixed order)
  int color;           // corner reverse in det mode: the corner parity class run
};

enum AdjKernel {
  AK_FOLD_OUT = 0,  // adjoint of set_bvals on aout (ghosts folded into their sources)
  AK_FOLD_IN,       // ... on ain
  AK_ZERO,          // clear ain and the accumulated adjoint planes
  AK_CLEAR,         // clear all adjoint planes (start of a reverse pass)
  AK_UPDATE,        // Steps 9, 11a, 11b, 13
  AK_CORNER_C,      // Step 9 corner emf
  AK_RIEMANN_C1,    // Step 8b (corrector x1 fluxes)
  AK_RIEMANN_C2,    // Step 8c (corrector x2 fluxes)
  AK_HALF,          // Step 7 (d, M1, M2 and emf3 at t^{n+1/2})
  AK_TRANSVERSE,    // Steps 5, 6 and the half-step face fields of Step 4
  AK_CORNER_P,      // Step 4 corner emf
  AK_EMF_CC,        // Step 3
  AK_PRED_X1,       // Steps 1c-1e (source, prim->cons, predictor flux) per x1 face
  AK_PLM_X1,        // Step 1b (PLM) per cell
  AK_PLMG_X1,       // Step 1a (cons->prim) per cell, gathering the PLM adjoints
  AK_PRED_X2,       // Steps 2c-2e per x2 face
  AK_PLM_X2,        // Step 2b per cell
  AK_PLMG_X2,       // Step 2a per cell
  AK_FOLD_OUT_DET,  // AK_FOLD_OUT as an ordered gather per source cell
  AK_COUNT
};

}  // namespace athmi

extern "C" hipError_t athmi_adjoint_launch(const athmi::AdjArgs* a, int kernel, hipStream_t s);
