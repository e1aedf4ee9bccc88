// adjoint_capi.cpp -- host runtime of the differentiation form (include/
// athmi_adjoint.h): a fixed-dt rollout of N CTU+CT steps of one instance
// (set_bvals, then N times integrate_2d + set_bvals) with a checkpoint every
// K steps, and its reverse: per segment of K steps the steps are re-run from
// the checkpoint, each step reading its input state from and writing its
// output state and forward scratch (predictor face states, conserved and
// primitive, and fluxes, predictor corner emf, corrector fluxes) into the
// segment's per-step slots; then each step is reversed by the adjoint
// kernels, last step first. With two banks of slots (when they fit in device
// memory) the re-run of segment s-1 on the handle's stream overlaps the
// reversal of segment s on a second stream. The forward pass runs its last
// segment through the slots itself, so the reverse starts without a re-run
// (with one segment, K >= N, it re-runs nothing).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "adjoint_kernels.hpp"
#include "athmi_adjoint.h"
#include "runtime.hpp"

namespace {

// forward scratch of one step: predictor, predictor corner emf, corrector fluxes
const int kSeqForwardScratch[] = {SK_BEGIN_ONCE, SK_PRED_X1_SAVE, SK_PRED_X2_SAVE, SK_CORNER_P, SK_CORR,
                                  SK_CORR2};
// the rest of integrate_2d after kSeqForwardScratch, and the set_bvals that
// follows it. integrate_2d leaves the ghost cells as they were (the drop-in's
// carry copies them into the new buffer), but the set_bvals right after it
// writes every ghost cell from non-ghost ones when Nx1, Nx2 >= 4, which the
// adjoint requires: its steps skip the carry
const int kSeqFinishStep[] = {SK_CORNER_C, SK_FINAL, SK_END_ONCE};
const int kSeqGhostCur[] = {SK_GHOST_CUR};
// a whole step so: integrate_2d without the carry (then set_bvals)
const int kSeqRolloutStep[] = {SK_BEGIN_ONCE, SK_PRED_X1, SK_PRED_X2, SK_CORNER_P, SK_CORR,
                               SK_CORR2,      SK_CORNER_C, SK_FINAL,  SK_END_ONCE};
// the reverse of one step on two streams: the x1 and x2 halves of Steps 8
// and 1-2 are independent (their adjoints meet in ain through atomic adds)
const int kRevHead[] = {AK_FOLD_OUT, AK_ZERO, AK_UPDATE, AK_CORNER_C};  // then fork
const int kRevMainA[] = {AK_RIEMANN_C1, AK_HALF};                       // || AK_RIEMANN_C2
const int kRevMid[] = {AK_TRANSVERSE, AK_CORNER_P, AK_EMF_CC};          // after the join
const int kRevTailA[] = {AK_PRED_X1, AK_PLM_X1, AK_PLMG_X1};            // || x2 tail
const int kRevTailB[] = {AK_PRED_X2, AK_PLM_X2, AK_PLMG_X2};
// a slot of the segment's forward scratch: the step's NSCRATCH planes, a copy
// of its predictor corner emf (used at fusion depth 0, where the corrector's
// corner kernel overwrites S_EMF) and the predictors' NPRIMS primitive face
// states
constexpr int kSlotEmf = NSCRATCH;
constexpr int kSlotPrims = NSCRATCH + 1;
constexpr int kSlotPlanes = NSCRATCH + 1 + NPRIMS;

}  // namespace

struct athmi_adjoint {
  athmi_grid* g = nullptr;
  double dt = 0.0;
  int nsteps = 0, every = 0, nck = 0;
  int nbank = 1;            // banks of segment slots (2: re-run and reversal overlap)
  int cur = 0;              // host copy of the clock's current buffer
  bool have_forward = false;
  bool last_filled = false;  // the last segment's slots hold the forward pass's own steps
  bool det = false;         // deterministic reverse: one stream, ordered sums
  unsigned long long generation = 0;  // completed forward passes
  double* mem = nullptr;    // checkpoints | segment states | adjoint state x2 | adjoint planes
  double* ck = nullptr;
  double* seg = nullptr;     // per bank K + 1 states: the input of each step of a segment
  double* adj[2] = {nullptr, nullptr};
  double* ab = nullptr;
  double* segsc = nullptr;   // per bank K slots of kSlotPlanes: forward scratch of each step
  size_t state_doubles = 0;  // NSTATE * ps
  size_t slot_doubles = 0;   // kSlotPlanes * ps
  // device descriptors of the re-run: step m of a segment in bank b whose
  // first step reads state buffer p is dslot[(2 b + p) K + m] (state buffers
  // seg[b][m], seg[b][m+1], scratch slot [b][m])
  Inst* dslot = nullptr;
  hipStream_t astream = nullptr, bstream = nullptr;  // the reversal's two streams
  hipEvent_t ev_rerun[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr}, ev_sync = nullptr;
  hipEvent_t ev_fork1 = nullptr, ev_join1 = nullptr, ev_fork2 = nullptr, ev_join2 = nullptr;

  double* seg_state(int b, int m) const {
    return seg + ((size_t)b * (every + 1) + m) * state_doubles;
  }
  double* slot(int b, int m) const { return segsc + ((size_t)b * every + m) * slot_doubles; }
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
  int rc = adj_issue(h, kSeqRolloutStep, sizeof(kSeqRolloutStep) / sizeof(int));
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
static int dense_to_planes(athmi_adjoint* h, double* planes, const double* dense, hipStream_t s) {
  const Inst& I = h->g->hinst[0];
  for (int q = 0; q < NSTATE; q++)
    HIP_TRY(hipMemcpy2DAsync(planes + q * I.ps, I.pitch * sizeof(double),
                             dense + (size_t)q * I.sx * I.sy, I.sx * sizeof(double),
                             I.sx * sizeof(double), I.sy, hipMemcpyDeviceToDevice, s));
  return ATHMI_OK;
}
static int planes_to_dense(athmi_adjoint* h, double* dense, const double* planes, hipStream_t s) {
  const Inst& I = h->g->hinst[0];
  for (int q = 0; q < NSTATE; q++)
    HIP_TRY(hipMemcpy2DAsync(dense + (size_t)q * I.sx * I.sy, I.sx * sizeof(double),
                             planes + q * I.ps, I.pitch * sizeof(double), I.sx * sizeof(double),
                             I.sy, hipMemcpyDeviceToDevice, s));
  return ATHMI_OK;
}

// re-run segment s (steps n0..n1-1) from its checkpoint into bank b on the
// handle's stream. finish_last: also complete the segment's last step (its
// output state, after set_bvals, is then seg_state(b, n1 - n0)); the reverse
// does not need it.
static int rerun_segment(athmi_adjoint* h, int s, int b, bool finish_last = false) {
  const int n0 = s * h->every, n1 = std::min(n0 + h->every, h->nsteps);
  const size_t ps = h->g->hinst[0].ps;
  const Inst* slots = h->dslot + (size_t)(2 * b + h->cur) * h->every;
  int rc;
  if ((rc = copy_state(h, h->seg_state(b, 0), h->ck + (size_t)s * h->state_doubles))) return rc;
  for (int m = n0; m < n1; m++) {
    const Inst* ins = slots + (m - n0);
    double* sc = h->slot(b, m - n0);
    if ((rc = issue_on(h, ins, kSeqForwardScratch, sizeof(kSeqForwardScratch) / sizeof(int))))
      return rc;
    // at fusion depth 0 the corrector's corner kernel overwrites the
    // predictor corners; otherwise they stay in S_EMF and the reverse reads
    // them there
    if (!h->g->bl.shape.fuse_corner)
      HIP_TRY(hipMemcpyAsync(sc + kSlotEmf * ps, sc + S_EMF * ps, ps * sizeof(double),
                             hipMemcpyDeviceToDevice, h->g->stream));
    if (m + 1 < n1 || finish_last) {
      if ((rc = issue_on(h, ins, kSeqFinishStep, sizeof(kSeqFinishStep) / sizeof(int)))) return rc;
      h->cur ^= 1;  // k_end swaps the buffers: the output is seg[b][m + 1]
      if ((rc = issue_on(h, ins, kSeqGhostCur, 1))) return rc;
    }
  }
  return ATHMI_OK;
}

extern "C" {

int athmi_adjoint_create(const athmi_desc* desc, double dt, int nsteps, int every,
                         athmi_adjoint** out) {
  if (!desc || !out || nsteps <= 0 || !(dt > 0.0))
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
  const Inst& I = h->g->hinst[0];
  h->state_doubles = (size_t)NSTATE * I.ps;
  h->slot_doubles = (size_t)kSlotPlanes * I.ps;
  size_t free_b = 0, total_b = 0;
  const bool have_free = hipMemGetInfo(&free_b, &total_b) == hipSuccess;
  if (every <= 0) {
    // automatic: one segment (every step stored, no re-run) when its slots
    // leave a fifth of the free memory, else sqrt(N)
    const size_t one = h->state_doubles * (1 + (size_t)(nsteps + 1) + 2) + (size_t)NADJ * I.ps +
                       h->slot_doubles * (size_t)nsteps;
    every = (have_free && one * sizeof(double) < free_b / 5 * 4)
                ? nsteps
                : std::max(1, (int)std::lround(std::sqrt((double)nsteps)));
  }
  h->every = std::min(every, nsteps);
  h->nck = (nsteps + h->every - 1) / h->every;
  // two banks of segment slots when they leave a fifth of the free memory
  auto total_for = [&](int nb) {
    return h->state_doubles * (h->nck + (size_t)nb * (h->every + 1) + 2) + (size_t)NADJ * I.ps +
           h->slot_doubles * nb * h->every;
  };
  h->nbank = (h->nck > 1 && have_free && total_for(2) * sizeof(double) < free_b / 5 * 4) ? 2 : 1;
  if (hipMalloc(&h->mem, total_for(h->nbank) * sizeof(double)) != hipSuccess ||
      hipMalloc(&h->dslot, sizeof(Inst) * 4 * h->every) != hipSuccess ||
      hipStreamCreateWithFlags(&h->astream, hipStreamNonBlocking) != hipSuccess) {
    athmi_adjoint_destroy(h);
    return fail(ATHMI_ENOMEM, "athmi_adjoint_create: no room for checkpoints");
  }
  for (int b = 0; b < 2; b++)
    if (hipEventCreateWithFlags(&h->ev_rerun[b], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_rev[b], hipEventDisableTiming) != hipSuccess) {
      athmi_adjoint_destroy(h);
      return fail(ATHMI_EHIP, "athmi_adjoint_create: event creation failed");
    }
  if (hipStreamCreateWithFlags(&h->bstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_sync, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fork1, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join1, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fork2, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join2, hipEventDisableTiming) != hipSuccess) {
    athmi_adjoint_destroy(h);
    return fail(ATHMI_EHIP, "athmi_adjoint_create: event creation failed");
  }
  h->ck = h->mem;
  h->seg = h->ck + h->state_doubles * h->nck;
  h->adj[0] = h->seg + h->state_doubles * h->nbank * (h->every + 1);
  h->adj[1] = h->adj[0] + h->state_doubles;
  h->ab = h->adj[1] + h->state_doubles;
  h->segsc = h->ab + (size_t)NADJ * I.ps;
  std::vector<Inst> slots(4 * h->every, I);
  for (int b = 0; b < h->nbank; b++)
    for (int p = 0; p < 2; p++)
      for (int m = 0; m < h->every; m++) {
        Inst& S = slots[(2 * b + p) * h->every + m];
        const int cur = p ^ (m & 1);
        S.state[cur] = h->seg_state(b, m);
        S.state[cur ^ 1] = h->seg_state(b, m + 1);
        S.scratch = h->slot(b, m);
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
  if (h->astream) (void)hipStreamSynchronize(h->astream);
  if (h->mem) (void)hipFree(h->mem);
  if (h->dslot) (void)hipFree(h->dslot);
  for (int b = 0; b < 2; b++) {
    if (h->ev_rerun[b]) (void)hipEventDestroy(h->ev_rerun[b]);
    if (h->ev_rev[b]) (void)hipEventDestroy(h->ev_rev[b]);
  }
  for (hipEvent_t e : {h->ev_sync, h->ev_fork1, h->ev_join1, h->ev_fork2, h->ev_join2})
    if (e) (void)hipEventDestroy(e);
  if (h->bstream) (void)hipStreamSynchronize(h->bstream);
  if (h->astream) (void)hipStreamDestroy(h->astream);
  if (h->bstream) (void)hipStreamDestroy(h->bstream);
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
  h->have_forward = false;
  h->last_filled = false;
  if ((rc = dense_to_planes(h, live_state(h), state0, h->g->stream))) return rc;
  HIP_TRY(h->g->launch(&h->g->bl, SK_GHOST_CUR, h->g->stream));
  const int slast = h->nck - 1, blast = slast % h->nbank;
  for (int n = 0; n < slast * h->every; n++) {
    if (n % h->every == 0 &&
        (rc = copy_state(h, h->ck + (size_t)(n / h->every) * h->state_doubles, live_state(h))))
      return rc;
    if ((rc = rollout_step(h))) return rc;
  }
  // the last segment: its checkpoint, then its steps through the slots the
  // reverse reads (the same kernels as a rollout step, with the primitive
  // face states stored too)
  if ((rc = copy_state(h, h->ck + (size_t)slast * h->state_doubles, live_state(h)))) return rc;
  if ((rc = rerun_segment(h, slast, blast, true))) return rc;
  const int n_last = h->nsteps - slast * h->every;
  if ((rc = planes_to_dense(h, stateN, h->seg_state(blast, n_last), h->g->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(h->g->stream));
  h->have_forward = true;
  h->last_filled = true;
  h->generation++;
  return ATHMI_OK;
}

int athmi_adjoint_every(const athmi_adjoint* h, int* every) {
  if (!h || !every) return fail(ATHMI_EINVAL, "athmi_adjoint_every: null argument");
  *every = h->every;
  return ATHMI_OK;
}

int athmi_adjoint_set_deterministic(athmi_adjoint* h, int on) {
  if (!h) return fail(ATHMI_EINVAL, "null handle");
  h->det = on != 0;
  return ATHMI_OK;
}

int athmi_adjoint_generation(const athmi_adjoint* h, unsigned long long* gen) {
  if (!h || !gen) return fail(ATHMI_EINVAL, "athmi_adjoint_generation: null argument");
  *gen = h->generation;
  return ATHMI_OK;
}

int athmi_adjoint_backward(athmi_adjoint* h, const double* gradN, double* grad0) {
  if (!h || !gradN || !grad0) return fail(ATHMI_EINVAL, "athmi_adjoint_backward: null argument");
  if (!h->have_forward)
    return fail(ATHMI_ESTATE, "athmi_adjoint_backward: run athmi_adjoint_forward first");
  int rc;
  hipStream_t R = h->g->stream, A = h->astream, B = h->bstream;
  double* aout = h->adj[0];
  double* ain = h->adj[1];
  if ((rc = dense_to_planes(h, aout, gradN, R))) return rc;
  AdjArgs a;
  a.g = h->g->hinst[0];
  a.ab = h->ab;
  a.dt = h->dt;
  a.ain = ain;
  a.det = h->det ? 1 : 0;
  a.color = 0;
  // the deterministic reverse runs every kernel of a step on one stream, in
  // the order of the two-stream schedule, with the scatters replaced by
  // ordered gathers (fold, predictor sources) or disjoint launches (corners)
  const int fold = h->det ? AK_FOLD_OUT_DET : AK_FOLD_OUT;
  if (h->det) B = A;
  HIP_TRY(athmi_adjoint_launch(&a, AK_CLEAR, R));
  HIP_TRY(hipEventRecord(h->ev_sync, R));
  HIP_TRY(hipStreamWaitEvent(A, h->ev_sync, 0));
  HIP_TRY(hipEventRecord(h->ev_join2, A));  // nothing pending on B yet
  const size_t ps = a.g.ps;
  bool rev_pending[2] = {false, false};
  const int blast = (h->nck - 1) % h->nbank;
  if (!h->last_filled && (rc = rerun_segment(h, h->nck - 1, blast))) return rc;
  for (int s = h->nck - 1; s >= 0; s--) {
    const int b = s % h->nbank;
    const int n0 = s * h->every, n1 = std::min(n0 + h->every, h->nsteps);
    // reverse segment s on the reversal stream once its re-run is done
    HIP_TRY(hipEventRecord(h->ev_rerun[b], R));
    HIP_TRY(hipStreamWaitEvent(A, h->ev_rerun[b], 0));
    for (int m = n1 - 1; m >= n0; m--) {
      const double* sc = h->slot(b, m - n0);
      a.st = h->seg_state(b, m - n0);
      a.sc = sc;
      a.emfp = sc + (h->g->bl.shape.fuse_corner ? (int)S_EMF : kSlotEmf) * ps;
      a.pw = sc + kSlotPrims * ps;
      a.ain = ain;
      a.aout = aout;
      // the previous step's x2 tail on B wrote ain planes this step folds
      HIP_TRY(hipStreamWaitEvent(A, h->ev_join2, 0));
      for (int k : kRevHead) HIP_TRY(athmi_adjoint_launch(&a, k == AK_FOLD_OUT ? fold : k, A));
      HIP_TRY(hipEventRecord(h->ev_fork1, A));
      HIP_TRY(hipStreamWaitEvent(B, h->ev_fork1, 0));
      HIP_TRY(athmi_adjoint_launch(&a, AK_RIEMANN_C2, B));
      HIP_TRY(hipEventRecord(h->ev_join1, B));
      for (int k : kRevMainA) HIP_TRY(athmi_adjoint_launch(&a, k, A));
      HIP_TRY(hipStreamWaitEvent(A, h->ev_join1, 0));
      for (int k : kRevMid) HIP_TRY(athmi_adjoint_launch(&a, k, A));
      HIP_TRY(hipEventRecord(h->ev_fork2, A));
      HIP_TRY(hipStreamWaitEvent(B, h->ev_fork2, 0));
      for (int k : kRevTailB) HIP_TRY(athmi_adjoint_launch(&a, k, B));
      HIP_TRY(hipEventRecord(h->ev_join2, B));
      for (int k : kRevTailA) HIP_TRY(athmi_adjoint_launch(&a, k, A));
      std::swap(ain, aout);
    }
    HIP_TRY(hipStreamWaitEvent(A, h->ev_join2, 0));
    HIP_TRY(hipEventRecord(h->ev_rev[b], A));
    rev_pending[b] = true;
    // meanwhile re-run segment s-1 into its bank, once that bank's last
    // reversal is done
    if (s > 0) {
      const int bn = (s - 1) % h->nbank;
      if (rev_pending[bn]) HIP_TRY(hipStreamWaitEvent(R, h->ev_rev[bn], 0));
      if (bn == blast) h->last_filled = false;  // the last segment's slots are reused
      if ((rc = rerun_segment(h, s - 1, bn))) return rc;
    }
  }
  // the set_bvals that precedes the first step
  a.aout = aout;
  HIP_TRY(athmi_adjoint_launch(&a, fold, A));
  if ((rc = planes_to_dense(h, grad0, aout, A))) return rc;
  HIP_TRY(hipStreamSynchronize(A));
  HIP_TRY(hipStreamSynchronize(R));
  return ATHMI_OK;
}

}  // extern "C"
