/*
 * athmi_adjoint.h -- differentiation form of the 2-D CTU + CT MHD step:
 * hand-written adjoint (reverse-mode) kernels of the whole step chained over
 * N fixed-dt steps with sqrt(N) checkpointing, plus host evaluations of the
 * per-routine adjoints. No reference counterpart: the reference has no
 * derivatives; each routine below is the adjoint of the reference routine
 * named beside it.
 */
#ifndef ATHMI_ADJOINT_H
#define ATHMI_ADJOINT_H

#include "athmi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A rollout of one instance: set_bvals, then nsteps times integrate_2d with
 * the fixed dt followed by set_bvals (main.c:221-318 with new_dt bypassed).
 * The state is held as dense device arrays [NSTATE=11][Nx2+8][Nx1+8] in the
 * order d, M1, M2, M3, E, B1c, B2c, B3c, B1i, B2i, B3i. The forward pass keeps
 * a checkpoint every `every` steps; the backward pass recomputes each segment
 * from its checkpoint and reverses it step by step. The forward pass stores
 * the last segment's steps as it runs them, so the backward pass re-runs all
 * segments but that one; with every >= nsteps (one segment, each step's 99
 * planes held: about 56 GB at 512^2, N = 256) it re-runs nothing. every <= 0
 * chooses: one segment when its slots leave a fifth of the free device
 * memory, else round(sqrt(nsteps)) (athmi_adjoint_every gives the choice).
 * Requires ATHMI_EXACT
 * arithmetic (the adjoint is linearised at the branches of that forward
 * pass) and Nx1, Nx2 >= 4. Ghost values of state0 are ignored (set_bvals
 * overwrites them), so their gradient is zero. */
typedef struct athmi_adjoint athmi_adjoint;
int athmi_adjoint_create(const athmi_desc *desc, double dt, int nsteps, int every,
                         athmi_adjoint **out);
int athmi_adjoint_destroy(athmi_adjoint *h);
/* the checkpoint spacing in use (every <= 0 at create: the automatic choice);
 * no reference counterpart */
int athmi_adjoint_every(const athmi_adjoint *h, int *every);
int athmi_adjoint_set_stream(athmi_adjoint *h, void *hip_stream);
/* state0, stateN: device pointers (dense layout above). Both calls run on the
 * handle's stream and synchronise it before returning. */
int athmi_adjoint_forward(athmi_adjoint *h, const double *state0, double *stateN);
/* gradN = dL/dstateN, grad0 = dL/dstate0 (device pointers, dense layout);
 * valid after athmi_adjoint_forward with the same state0 */
int athmi_adjoint_backward(athmi_adjoint *h, const double *gradN, double *grad0);
/* on = 1: the reverse pass sums every adjoint in a fixed order, so two
 * backward passes of one forward pass give bitwise equal gradients. It runs
 * on one stream, folds the ghost adjoints and the predictor-face source
 * terms as ordered gathers and reverses the corner emf in four disjoint
 * launches; on = 0 (the default) uses two streams and FP64 atomic adds,
 * whose order varies at roundoff level. No reference counterpart (the
 * reference has no derivatives). */
int athmi_adjoint_set_deterministic(athmi_adjoint *h, int on);
/* Number of completed athmi_adjoint_forward calls on the handle. The
 * checkpoints belong to the latest forward pass only: a caller that records
 * the generation after its forward pass and compares it before its backward
 * pass detects that another forward pass on the same handle came between. */
int athmi_adjoint_generation(const athmi_adjoint *h, unsigned long long *gen);

/* Per-routine vector-Jacobian products on the host (same source as the
 * adjoint kernels). Vectors use the Cons1D / Prim1D member order
 * (athena.h:75-97); eigenmatrices are row-major 7x7. */
/* convert_var.c:37-55 */
void athmi_vjp_cons_to_prim(double gamma, const double *u, double bx, const double *wb,
                            double *ub, double *bxb);
/* convert_var.c:64-92 */
void athmi_vjp_prim_to_cons(double gamma, const double *w, double bx, const double *ub,
                            double *wb, double *bxb);
/* esystem_prim.c:36-235; in = (d, v1, p, b1, b2, b3) */
void athmi_vjp_esys_prim(double gamma, const double *in, const double *Rb, const double *Lb,
                         const double *evb, double *inb);
/* esystem_roe.c:56-292; in = (d, v1, v2, v3, h, b1, b2, b3, x, y); vectors = 0:
 * the eigenvalues-only call */
void athmi_vjp_esys_roe(double gamma, const double *in, const double *Rb, const double *Lb,
                        const double *evb, double *inb, int vectors);
/* flux_roe.c:54-259 (kind 0, returns its HLLE-fallback flag) and
 * flux_hlle.c:60-199 (kind 1) */
int athmi_vjp_flux(double gamma, int kind, const double *ul, const double *ur, const double *wl,
                   const double *wr, double bx, const double *fb, double *ulb, double *urb,
                   double *wlb, double *wrb, double *bxb);
/* lr_states_plm.c:64-248 for one cell i: out_l = Wl[i+1], out_r = Wr[i] */
void athmi_vjp_plm(double gamma, const double *wm, const double *w0, const double *wp, double bxc,
                   double dtodx, const double *out_lb, const double *out_rb, double *out_l,
                   double *out_r, double *wmb, double *w0b, double *wpb, double *bxcb);

#ifdef __cplusplus
}
#endif

#endif /* ATHMI_ADJOINT_H */
