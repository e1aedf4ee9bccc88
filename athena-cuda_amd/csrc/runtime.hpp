// runtime.hpp -- internal: the handle behind athmi_grid (device state of the
// instances, tile work-list, stream and step graph) shared by the C-ABI
// translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "athmi.h"
#include "host_problems.hpp"
#include "step_kernels.hpp"

ATHMI_DECLARE_VARIANT(exact)
ATHMI_DECLARE_VARIANT(fast)

using namespace athmi;

namespace athmi_rt {

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(ATHMI_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

typedef hipError_t (*LaunchFn)(const void*, int, hipStream_t);

// transfer.hip: caller layout (gas[sy*sx][8] | b1i | b2i | b3i in `stage`)
// <-> the SoA planes of one state buffer
}  // namespace athmi_rt
namespace athmi {
hipError_t launch_state_unpack(const double* stage, double* st, size_t ps, int pitch, int sx,
                               int sy, hipStream_t s);
hipError_t launch_state_pack(const double* st, double* stage, size_t ps, int pitch, int sx,
                             int sy, hipStream_t s);
}  // namespace athmi
namespace athmi_rt {

// kernel sequences of the reference's routines
const int kSeqStep[] = {SK_BEGIN_STEP, SK_PRED_X1, SK_PRED_X2, SK_CORNER_P,   SK_CORR,
                        SK_CORR2,      SK_CORNER_C, SK_FINAL,  SK_GHOST_NEXT, SK_END_STEP};
const int kSeqIntegrate[] = {SK_BEGIN_ONCE, SK_PRED_X1,  SK_PRED_X2, SK_CORNER_P,    SK_CORR,
                             SK_CORR2,      SK_CORNER_C, SK_FINAL,   SK_GHOST_CARRY, SK_END_ONCE};
const int kSeqNewDt[] = {SK_CLEAR, SK_DTI, SK_DT_FINAL};
const int kSeqDivb[] = {SK_CLEAR, SK_DIVB};
const char* const kStepNames[] = {"k_begin",   "k_pred_x1",      "k_pred_x2", "k_corner<pred>",
                                  "k_corr<1>", "k_corr<2>",      "k_corner<corr>", "k_final",
                                  "k_ghost",   "k_end"};
constexpr int kStepLen = sizeof(kSeqStep) / sizeof(int);

inline size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

}  // namespace athmi_rt

using namespace athmi_rt;

struct athmi_grid {
  int n = 0;
  int variant = ATHMI_EXACT;
  std::vector<athmi_desc> desc;
  std::vector<Inst> hinst;
  Inst* dinst = nullptr;
  Clock* dclk = nullptr;
  double* dmem = nullptr;
  size_t dmem_bytes = 0;
  BatchLaunch bl{};
  LaunchFn launch = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  std::vector<std::vector<double>> cpaw_soln;  // initial cpaw2d states (8 planes)
  void* outbuf = nullptr;                      // device staging of the output formats
  size_t outbuf_bytes = 0;

  void drop_graph() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
  int refresh_tiles() {
    int tot[NKINDS] = {0};
    for (int k = 0; k < n; k++) {
      Inst& I = hinst[k];
      for (int kind = 0; kind < NKINDS; kind++) {
        I.tile0[kind] = tot[kind];
        tot[kind] += tiles_of(kind, I.nx1, I.nx2, bl.shape.x2_seg, bl.shape.corr_rows,
                              bl.shape.final_rows);
      }
    }
    for (int kind = 0; kind < NKINDS; kind++) bl.ntiles[kind] = tot[kind];
    bl.flux_mask = 0;
    bl.stale_ghosts = 0;
    for (int k = 0; k < n; k++) {
      bl.flux_mask |= 1 << hinst[k].flux;
      if (ghost_fill_reads_stale(hinst[k])) bl.stale_ghosts = 1;
    }
    HIP_TRY(hipMemcpy(dinst, hinst.data(), sizeof(Inst) * n, hipMemcpyHostToDevice));
    drop_graph();
    return ATHMI_OK;
  }
  int read_clock(int k, Clock& c) {
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipMemcpy(&c, dclk + k, offsetof(Clock, divb_bits) + sizeof(unsigned long long),
                      hipMemcpyDeviceToHost));
    return ATHMI_OK;
  }
  int write_clock_head(int k, const Clock& c) {
    HIP_TRY(hipMemcpy(dclk + k, &c, offsetof(Clock, divb_bits), hipMemcpyHostToDevice));
    return ATHMI_OK;
  }
};

// the reference's Grid geometry of one instance (init_grid.c:25-118)
inline HostGeom geom_of(const athmi_grid* g, int inst) {
  const athmi_desc& d = g->desc[inst];
  const Inst& I = g->hinst[inst];
  HostGeom G;
  G.Nx1 = d.nx1;
  G.Nx2 = d.nx2;
  G.is = I.is;
  G.ie = I.ie;
  G.js = I.js;
  G.je = I.je;
  G.sx = I.sx;
  G.sy = I.sy;
  G.dx1 = I.dx1;
  G.dx2 = I.dx2;
  G.x1_0 = d.x1min;
  G.x2_0 = d.x2min;
  G.idisp = -I.is;
  G.jdisp = -I.js;
  G.gamma_1 = I.g1;
  return G;
}

inline int issue_seq(athmi_grid* g, const int* seq, size_t n) {
  for (size_t k = 0; k < n; k++) HIP_TRY(g->launch(&g->bl, seq[k], g->stream));
  return ATHMI_OK;
}
#define issue(g, seq) issue_seq((g), (seq), sizeof(seq) / sizeof((seq)[0]))

