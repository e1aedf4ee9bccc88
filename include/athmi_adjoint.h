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

#ifdef __cplusplus
extern "C" {
#endif

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
