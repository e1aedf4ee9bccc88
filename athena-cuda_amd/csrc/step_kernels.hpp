// step_kernels.hpp -- data layout shared by the host side and the HIP kernels
// of the CTU+CT step. One "instance" is one 2-D grid (the reference's Grid,
// athena.h:106-120); the plain step is a batch of one instance.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace athmi {

// field planes of a state buffer, in Gas order (athena.h:43-52) + interface B
enum Field { F_D = 0, F_M1, F_M2, F_M3, F_E, F_B1C, F_B2C, F_B3C, F_B1I, F_B2I, F_B3I, NSTATE = 11 };

// scratch planes (each Cons1D-shaped group is 7 planes d,Mx,My,Mz,E,By,Bz)
enum Scratch {
  S_UL1 = 0,    // x1-face left states (conserved), predictor
  S_UR1 = 7,    // x1-face right states
  S_F1 = 14,    // x1 predictor fluxes
  S_UL2 = 21,   // x2-face left states (x2 sweep order d,M2,M3,M1,E,B3c,B1c)
  S_UR2 = 28,
  S_F2 = 35,    // x2 predictor fluxes
  S_F1C = 42,   // x1 corrector fluxes
  S_F2C = 49,   // x2 corrector fluxes
  S_ECC = 56,   // cell-centred emf3 at t^n
  S_EMF = 57,   // corner emf3 (predictor, then corrector)
  S_ECH = 58,   // cell-centred emf3 at t^{n+1/2}
  NSCRATCH = 59
};

// optional planes of the predictor's primitive face states, after the 1/2 dt
// MHD source (the Wl, Wr that Steps 1e / 2e hand to the Riemann solver)
enum PrimSave { P_WL1 = 0, P_WR1 = 7, P_WL2 = 14, P_WR2 = 21, NPRIMS = 28 };

enum BcKind { BC_REFLECT = 1, BC_OUTFLOW = 2, BC_PERIODIC = 4 };

constexpr int kDtSlots = 256;  // partial max slots of the fused dt reduction

// Failure bits of an instance, raised by the step on the state it produces
// (the reference only floors, convert_var.c:49 and new_dt.c:54-55, and steps
// on). Once any is set the main-loop step no longer runs on that instance.
enum FailBits {
  FAIL_NONFINITE = 1,  // a NaN or infinity in d, M1..M3, E or the cell-centred B
  FAIL_DENSITY = 2,    // d <= 0
  FAIL_PRESSURE = 4,   // Gamma_1 (E - KE - ME) below TINY_NUMBER: the pressure floor engages
};

// Device-resident clock of one instance (Grid.dt/time/nstep of athena.h:113-114
// plus the main-loop limits of main.c:217-218).
struct Clock {
  double dt, time, tlim, cour_no;
  double tpause;  // the main-loop step does not run once time >= tpause (an output is due)
  int nstep, nlim, active, cur;  // cur: which ping-pong state buffer holds the state
  int fail;        // FailBits of the states produced so far (0: healthy)
  int fail_nstep;  // nstep of the first state that raised a bit (-1: none)
  int fail_stop;   // FailBits that stop the main-loop step (the others are only reported)
  int steps_run;
  unsigned long long divb_bits;           // max |div B| (bit pattern), diagnostics
  unsigned long long dti_slot[kDtSlots];  // bit patterns of non-negative doubles
};

struct Inst {
  int nx1, nx2, is, ie, js, je, sx, sy, pitch, flux;
  int bc[4];          // ix1, ox1, ix2, ox2
  double dx1, dx2, gamma, g1, g2;
  double* state[2];   // two ping-pong buffers of NSTATE planes
  double* scratch;    // NSCRATCH planes
  double* prims;      // NPRIMS planes written by SK_PRED_X1_SAVE / SK_PRED_X2_SAVE
  size_t ps;          // plane stride in doubles
  Clock* clk;
  // first global tile of this instance for each kernel kind
  int tile0[8];
};

// Tile kinds (one per kernel family); tile counts per instance come from
// tiles_of() below so host and device agree on the decomposition.
enum KernelKind {
  K_PRED_X1 = 0,  // rows js-2..je+2, faces is-1..ie+2 in segments of kX1Faces
  K_PRED_X2,      // columns is-2..ie+2 in chunks of kX2Block from column 0 (row segments
                  // start on a 128-byte line: the pitch is a multiple of 16), faces
                  // js-1..je+2 in segments
  K_CORR,         // cells [js-1,je+1] x [is-1,ie+1]
  K_CORNER_P,     // corners [js-1,je+2] x [is-1,ie+2]
  K_CORNER_C,     // corners [js,je+1] x [is,ie+1]
  K_FINAL,        // faces/cells [js,je+1] x [is,ie+1]
  K_GHOST,        // ghost cells of all planes
  K_INTERIOR,     // cells [js,je] x [is,ie]
  NKINDS
};

// Work decomposition (tile sizes) shared by host and device
// threads of a predictor-x1 tile (one row segment): two waves, so that the
// workgroups resident on a CU pass their block barriers independently (at one
// wave per SIMD, 4096^2: 2.56 ms with 256 threads, 2.45 with 64, 2.40 with
// 128; at two waves per SIMD the three are within 1%)
constexpr int kX1Block = 128;
constexpr int kX1Faces = kX1Block - 3;      // faces produced per x1 tile
constexpr int kX2Block = 64;                // columns of a predictor-x2 tile
constexpr int kTileX = 64, kTileY = 4;      // 2-D tiles of the cell/corner kernels
constexpr int kGhostBlock = 256;

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// rows of a final-update tile (64 x 4, 8 or 16 cells; anything else: 4)
inline int final_rows_of(int rows) { return (rows == 8 || rows == 16) ? rows : kTileY; }

// number of tiles of kind k for a grid of nx1 x nx2 zones (x2seg: faces
// marched per thread by the x2 predictor; corr_rows, final_rows: tile heights
// of the corrector and the final update)
// Default faces marched per x2-predictor lane for `cols` lane-columns of x2
// faces (64-column chunks x rows over the batch): shorter segments give more
// workgroups (less tail at the end of the launch) for more halo rows per face.
// The measured best depth grows about as sqrt(cols)/50: 1 at 256^2, 3 at
// 512^2-1024^2, 5 at 2048^2, 8-24 at 4096^2, all within 2-3% of each other.
inline int default_x2_seg(long cols) {
  const long s = std::lround(std::sqrt((double)cols) / 50.0);
  return (int)(s < 1 ? 1 : (s > 16 ? 16 : s));
}

inline int tiles_of(int k, int nx1, int nx2, int x2seg, int corr_rows, int final_rows) {
  const int sx = nx1 + 8;
  switch (k) {
    case K_PRED_X1: return (nx2 + 4) * cdiv(nx1 + 4, kX1Faces);
    case K_PRED_X2: return cdiv(nx1 + 6, kX2Block) * cdiv(nx2 + 4, x2seg);
    case K_CORR: return cdiv(nx1 + 3, kTileX) * cdiv(nx2 + 3, corr_rows);
    case K_CORNER_P: return cdiv(nx1 + 4, kTileX) * cdiv(nx2 + 4, kTileY);
    case K_CORNER_C: return cdiv(nx1 + 2, kTileX) * cdiv(nx2 + 2, kTileY);
    case K_FINAL: return cdiv(nx1 + 1, kTileX) * cdiv(nx2 + 1, final_rows_of(final_rows));
    case K_GHOST: return cdiv(8 * sx + (nx2 + 1) * 8, kGhostBlock);
    default: return cdiv(nx1, kTileX) * cdiv(nx2, kTileY);
  }
}

// whether an instance's set_bvals reads ghosts as they were before it (an
// inner reflecting fill across fewer cells than the four ghost layers copies
// outer ghosts the outer fill has not reached yet, set_bvals.c:312-366)
inline bool ghost_fill_reads_stale(const Inst& g) {
  return (g.bc[0] == BC_REFLECT && g.nx1 < 4) || (g.bc[2] == BC_REFLECT && g.nx2 < 4);
}

// Launch shape chosen by the autotuner
struct LaunchShape {
  int x2_seg;        // faces marched per thread in the x2 predictor
  int persist;       // 0: one tile per workgroup; >0: workgroups per CU of a persistent grid
  int xcd_remap;     // 1: contiguous runs of tiles per XCD (L2 locality), 0: launch order
  int corr_rows;     // rows of a corrector tile (kTileX x corr_rows faces): 1, 2 or 4
  int x2_kind;       // x2 predictor: 0 = stencil marched in registers (one wave per
                     // SIMD), 1 = stencil marched in LDS slots (two waves per SIMD)
  int fuse_corner;   // fusion depth. 1: the Step 9 corner emf is computed inside the
                     // final kernel's tiles (no k_corner<corr> launch, no corner plane
                     // round trip); 2: also the Step 4 corner emf inside the x1
                     // corrector's tiles (no k_corner<pred> launch)
  int strip;         // tile order of the corrector and final kernels: 0 = row by row,
                     // s > 0 = strips of s tile columns walked row by row
  int final_rows;    // rows of a final-update tile: 4, 8 or 16 (final_rows_of)
};

// Everything a launcher needs for one batch (device descriptor array, tile
// totals per kernel kind, the tuned shape and the CU count)
struct BatchLaunch {
  const Inst* insts;
  int ninst;
  int ntiles[NKINDS];
  int stale_ghosts;  // some instance's ghost fill reads stale ghosts (ghost_map.hpp)
  LaunchShape shape;
  int ncu;
  int flux_mask;  // bit f set: some instance uses Riemann solver f (kFluxRoe, kFluxHlle)
};

// Kernels of the pipeline, launched one at a time by athmi_<variant>_launch
enum StepKernel {
  SK_BEGIN_STEP = 0,  // k_begin, main-loop mode
  SK_BEGIN_ONCE,      // k_begin, unconditional (drop-in integrate_2d)
  SK_PRED_X1,
  SK_PRED_X2,
  SK_CORNER_P,
  SK_CORR,            // corrector, x1 faces
  SK_CORR2,           // corrector, x2 faces and emf3 at t^{n+1/2}
  SK_CORNER_C,
  SK_FINAL,
  SK_GHOST_CUR,       // set_bvals on the current buffer
  SK_GHOST_NEXT,      // set_bvals on the buffer this step produces (respects the clock)
  SK_GHOST_CARRY,     // carry current ghost values into the produced buffer
  SK_END_STEP,        // k_end, main-loop mode
  SK_END_ONCE,        // k_end, swap only
  SK_CLEAR,           // reset reduction slots
  SK_DTI,             // new_dt partial maxima
  SK_DT_FINAL,        // new_dt.c:93-97
  SK_DIVB,
  SK_PRED_X1_SAVE,    // predictors that also store their primitive face states
  SK_PRED_X2_SAVE,    // (Inst::prims), for the adjoint's re-run
  SK_COUNT
};

}  // namespace athmi

// Launcher exported by each arithmetic variant of step_kernels.hip
// (athmi_exact_launch built with -ffp-contract=off, athmi_fast_launch with FMA).
#define ATHMI_DECLARE_VARIANT(V) \
  extern "C" hipError_t athmi_##V##_launch(const void* bl, int kernel, hipStream_t s);
