// adjoint_capi.cpp -- host runtime of the differentiation form (include/
// athmi_adjoint.h): a fixed-dt rollout of N CTU+CT steps of one instance
// (set_bvals, then N times integrate_2d + set_bvals) with a checkpoint every
// K steps, and its reverse: per segment of K steps the steps are re-run from
// the checkpoint, each step reading its input state from and writing its
// output state and forward scratch (predictor face states, conserved and
// primitive, and fluxes, predictor corner emf, corrector fluxes) into the
// segment's per-step slots; then each step is reversed by the adjoint
// kernels, last step first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "adjoint_kernels.hpp"
#include "athmi_adjoint.h"
#include "runtime.hpp"

namespace {

// forward scratch of one step: predictor, predictor corner emf, corrector fluxes
const int kSeqForwardScratch[] = {SK_BEGIN_ONCE, SK_PRED_X1, SK_PRED_X2, SK_CORNER_P, SK_CORR,
                                  SK_CORR2};
// the rest of integrate_2d after kSeqForwardScratch
const int kSeqFinishStep[] = {SK_CORNER_C, SK_FINAL, SK_GHOST_CARRY, SK_END_ONCE};
const int kSeqGhostCur[] = {SK_GHOST_CUR};
const int kSeqReverse[] = {AK_FOLD_OUT,  AK_ZERO,       AK_UPDATE,   AK_CORNER_C, AK_RIEMANN_C,
                           AK_HALF,      AK_TRANSVERSE, AK_CORNER_P, AK_EMF_CC,   AK_PRED_X1,
                           AK_PLM_X1,    AK_PRED_X2,    AK_PLM_X2};
// a slot of the segment's forward scratch: the step's NSCRATCH planes, its
// predictor corner emf (the corrector's corner kernel overwrites S_EMF) and
// the predictors' NPRIMS primitive face states
constexpr int kSlotEmf = NSCRATCH;
constexpr int kSlotPrims = NSCRATCH + 1;
constexpr int kSlotPlanes = NSCRATCH + 1 + NPRIMS;

}  // namespace

struct athmi_adjoint {
  athmi_grid* g = nullptr;
  double dt = 0.0;
  int nsteps = 0, every = 0, nck = 0;
  int cur = 0;              // host copy of the clock's current buffer
  bool have_forward = false;
  double* mem = nullptr;    // checkpoints | segment states | adjoint state x2 | adjoint planes
  double* ck = nullptr;
  double* seg = nullptr;     // K + 1 states: the input of each step of a segment
  double* adj[2] = {nullptr, nullptr};
  double* ab = nullptr;
  double* segsc = nullptr;   // K slots of kSlotPlanes: forward scratch of each step
  size_t state_doubles = 0;  // NSTATE * ps
  size_t slot_doubles = 0;   // kSlotPlanes * ps
  // 2K device descriptors of the re-run: step m of a segment whose first step
  // reads state buffer p is dslot[p * K + m] (state buffers seg[m], seg[m+1],
  // scratch slot m)
  Inst* dslot = nullptr;
};

// launch a kernel sequence of the instance through the descriptor ins
static int issue_on(athmi_adjoint* h, const Inst* ins, const int* seq, size_t n) {
  BatchLaunch bl = h->g->bl;
  bl.insts = ins;
  for (size_t k = 0; k < n; k++) HIP_TRY(h->g->launch(&bl, seq[k], h->g->stream));
  return ATHMI_OK;
}

static int adj_issue(athmi_adjoint* h, const int* seq, size_t n) {
  return issue_seq(h->g, seq, n);
}

// one step of the rollout: integrate_2d with the clock's dt, then set_bvals
static int rollout_step(athmi_adjoint* h) {
  int rc = adj_issue(h, kSeqIntegrate, sizeof(kSeqIntegrate) / sizeof(int));
  if (rc) return rc;
  h->cur ^= 1;  // k_end swaps the buffers
  HIP_TRY(h->g->launch(&h->g->bl, SK_GHOST_CUR, h->g->stream));
  return ATHMI_OK;
}

static double* live_state(athmi_adjoint* h) { return h->g->hinst[0].state[h->cur]; }

static int copy_state(athmi_adjoint* h, double* dst, const double* src) {
  HIP_TRY(hipMemcpyAsync(dst, src, h->state_doubles * sizeof(double), hipMemcpyDeviceToDevice,
                         h->g->stream));
  return ATHMI_OK;
}

// dense [NSTATE][Nx2+8][Nx1+8] <-> pitched planes
static int dense_to_planes(athmi_adjoint* h, double* planes, const double* dense) {
  const Inst& I = h->g->hinst[0];
  for (int q = 0; q < NSTATE; q++)
    HIP_TRY(hipMemcpy2DAsync(planes + q * I.ps, I.pitch * sizeof(double),
                             dense + (size_t)q * I.sx * I.sy, I.sx * sizeof(double),
                             I.sx * sizeof(double), I.sy, hipMemcpyDeviceToDevice, h->g->stream));
  return ATHMI_OK;
}
static int planes_to_dense(athmi_adjoint* h, double* dense, const double* planes) {
  const Inst& I = h->g->hinst[0];
  for (int q = 0; q < NSTATE; q++)
    HIP_TRY(hipMemcpy2DAsync(dense + (size_t)q * I.sx * I.sy, I.sx * sizeof(double),
                             planes + q * I.ps, I.pitch * sizeof(double), I.sx * sizeof(double),
                             I.sy, hipMemcpyDeviceToDevice, h->g->stream));
  return ATHMI_OK;
}

extern "C" {

int athmi_adjoint_create(const athmi_desc* desc, double dt, int nsteps, int every,
                         athmi_adjoint** out) {
  if (!desc || !out || nsteps <= 0 || every <= 0 || !(dt > 0.0))
    return fail(ATHMI_EINVAL, "athmi_adjoint_create: bad arguments");
  if (desc->variant != ATHMI_EXACT)
    return fail(ATHMI_EINVAL,
                "athmi_adjoint_create: the adjoint linearises at the exact forward's branches; "
                "use ATHMI_EXACT");
  if (desc->nx1 < 4 || desc->nx2 < 4)
    return fail(ATHMI_EINVAL, "athmi_adjoint_create: Nx1, Nx2 >= 4 (ghost sources interior)");
  athmi_adjoint* h = new athmi_adjoint();
  int rc = athmi_create(desc, 1, &h->g);
  if (rc) {
    delete h;
    return rc;
  }
  h->dt = dt;
  h->nsteps = nsteps;
  h->every = std::min(every, nsteps);
  h->nck = (nsteps + h->every - 1) / h->every;
  const Inst& I = h->g->hinst[0];
  h->state_doubles = (size_t)NSTATE * I.ps;
  h->slot_doubles = (size_t)kSlotPlanes * I.ps;
  const size_t total = h->state_doubles * (h->nck + h->every + 3) + (size_t)NADJ * I.ps +
                       h->slot_doubles * h->every;
  if (hipMalloc(&h->mem, total * sizeof(double)) != hipSuccess ||
      hipMalloc(&h->dslot, sizeof(Inst) * 2 * h->every) != hipSuccess) {
    athmi_adjoint_destroy(h);
    return fail(ATHMI_ENOMEM, "athmi_adjoint_create: no room for checkpoints");
  }
  h->ck = h->mem;
  h->seg = h->ck + h->state_doubles * h->nck;
  h->adj[0] = h->seg + h->state_doubles * (h->every + 1);
  h->adj[1] = h->adj[0] + h->state_doubles;
  h->ab = h->adj[1] + h->state_doubles;
  h->segsc = h->ab + (size_t)NADJ * I.ps;
  std::vector<Inst> slots(2 * h->every, I);
  for (int p = 0; p < 2; p++)
    for (int m = 0; m < h->every; m++) {
      Inst& S = slots[p * h->every + m];
      const int cur = p ^ (m & 1);
      S.state[cur] = h->seg + (size_t)m * h->state_doubles;
      S.state[cur ^ 1] = h->seg + (size_t)(m + 1) * h->state_doubles;
      S.scratch = h->segsc + (size_t)m * h->slot_doubles;
      S.prims = S.scratch + (size_t)kSlotPrims * I.ps;
    }
  if (hipMemcpy(h->dslot, slots.data(), sizeof(Inst) * slots.size(), hipMemcpyHostToDevice) !=
      hipSuccess) {
    athmi_adjoint_destroy(h);
    return fail(ATHMI_EHIP, "athmi_adjoint_create: descriptor upload failed");
  }
  Clock c;
  if ((rc = h->g->read_clock(0, c)) || (rc = athmi_set_clock(h->g, 0, dt, 0.0, 0))) {
    athmi_adjoint_destroy(h);
    return rc;
  }
  h->cur = c.cur;
  *out = h;
  return ATHMI_OK;
}

int athmi_adjoint_destroy(athmi_adjoint* h) {
  if (!h) return ATHMI_OK;
  if (h->g && h->g->stream) (void)hipStreamSynchronize(h->g->stream);
  if (h->mem) (void)hipFree(h->mem);
  if (h->dslot) (void)hipFree(h->dslot);
  if (h->g) athmi_destroy(h->g);
  delete h;
  return ATHMI_OK;
}

int athmi_adjoint_set_stream(athmi_adjoint* h, void* hip_stream) {
  if (!h) return fail(ATHMI_EINVAL, "null handle");
  return athmi_set_stream(h->g, hip_stream);
}

int athmi_adjoint_forward(athmi_adjoint* h, const double* state0, double* stateN) {
  if (!h || !state0 || !stateN) return fail(ATHMI_EINVAL, "athmi_adjoint_forward: null argument");
  int rc;
  if ((rc = dense_to_planes(h, live_state(h), state0))) return rc;
  HIP_TRY(h->g->launch(&h->g->bl, SK_GHOST_CUR, h->g->stream));
  for (int n = 0; n < h->nsteps; n++) {
    if (n % h->every == 0 &&
        (rc = copy_state(h, h->ck + (size_t)(n / h->every) * h->state_doubles, live_state(h))))
      return rc;
    if ((rc = rollout_step(h))) return rc;
  }
  if ((rc = planes_to_dense(h, stateN, live_state(h)))) return rc;
  HIP_TRY(hipStreamSynchronize(h->g->stream));
  h->have_forward = true;
  return ATHMI_OK;
}

int athmi_adjoint_backward(athmi_adjoint* h, const double* gradN, double* grad0) {
  if (!h || !gradN || !grad0) return fail(ATHMI_EINVAL, "athmi_adjoint_backward: null argument");
  if (!h->have_forward)
    return fail(ATHMI_ESTATE, "athmi_adjoint_backward: run athmi_adjoint_forward first");
  int rc;
  double* aout = h->adj[0];
  double* ain = h->adj[1];
  if ((rc = dense_to_planes(h, aout, gradN))) return rc;
  AdjArgs a;
  a.g = h->g->hinst[0];
  a.ab = h->ab;
  a.dt = h->dt;
  a.ain = ain;
  HIP_TRY(athmi_adjoint_launch(&a, AK_CLEAR, h->g->stream));
  const size_t ps = a.g.ps;
  for (int s = h->nck - 1; s >= 0; s--) {
    const int n0 = s * h->every, n1 = std::min(n0 + h->every, h->nsteps);
    // re-run the segment from its checkpoint into the per-step slots
    const Inst* slot = h->dslot + (size_t)h->cur * h->every;
    if ((rc = copy_state(h, h->seg, h->ck + (size_t)s * h->state_doubles))) return rc;
    for (int m = n0; m < n1; m++) {
      const Inst* ins = slot + (m - n0);
      double* sc = h->segsc + (size_t)(m - n0) * h->slot_doubles;
      if ((rc = issue_on(h, ins, kSeqForwardScratch, sizeof(kSeqForwardScratch) / sizeof(int))))
        return rc;
      HIP_TRY(hipMemcpyAsync(sc + kSlotEmf * ps, sc + S_EMF * ps, ps * sizeof(double),
                             hipMemcpyDeviceToDevice, h->g->stream));
      if (m + 1 < n1) {
        if ((rc = issue_on(h, ins, kSeqFinishStep, sizeof(kSeqFinishStep) / sizeof(int))))
          return rc;
        h->cur ^= 1;  // k_end swaps the buffers: the output is seg[m + 1]
        if ((rc = issue_on(h, ins, kSeqGhostCur, 1))) return rc;
      }
    }
    for (int m = n1 - 1; m >= n0; m--) {
      const double* sc = h->segsc + (size_t)(m - n0) * h->slot_doubles;
      a.st = h->seg + (size_t)(m - n0) * h->state_doubles;
      a.sc = sc;
      a.emfp = sc + kSlotEmf * ps;
      a.pw = sc + kSlotPrims * ps;
      a.ain = ain;
      a.aout = aout;
      for (int k : kSeqReverse) HIP_TRY(athmi_adjoint_launch(&a, k, h->g->stream));
      std::swap(ain, aout);
    }
  }
  // the set_bvals that precedes the first step
  a.aout = aout;
  HIP_TRY(athmi_adjoint_launch(&a, AK_FOLD_OUT, h->g->stream));
  if ((rc = planes_to_dense(h, grad0, aout))) return rc;
  HIP_TRY(hipStreamSynchronize(h->g->stream));
  return ATHMI_OK;
}

}  // extern "C"
