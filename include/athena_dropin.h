/*
 * athena_dropin.h -- the reference's own CPU entry points for this path, with
 * their exact names and signatures, implemented on top of athmi.h. A program
 * written against the reference (src/main.c) links libathmi.so in place of
 * __integrate_2d__.o, new_dt.o, set_bvals.o and lr_states_plm.o and runs the
 * step on the GPU. Each call copies the Grid to the device and back, so it
 * keeps the reference's in-place semantics; a program that wants the state to
 * stay in HBM uses athmi_step instead.
 *
 * athmi_ref_grid has the memory layout of the reference's Grid
 * (athena.h:106-120, Real = double) so a Grid* can be passed as is.
 */
#ifndef ATHENA_DROPIN_H
#define ATHENA_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct athmi_ref_gas { double d, M1, M2, M3, E, B1c, B2c, B3c; } athmi_ref_gas;

typedef struct athmi_ref_grid {
  athmi_ref_gas **U;
  double **B1i, **B2i, **B3i;
  double x1_0, x2_0;
  double dx1, dx2;
  double dt, time;
  int nstep;
  int Nx1, Nx2;
  int is, ie;
  int js, je;
  int idisp, jdisp;
  char *outfilename;
} athmi_ref_grid;

/* prototypes.h:103-105 (integrate_2d.c) */
void integrate_init_2d(int Nx1, int Nx2);
void integrate_destruct_2d(void);
void integrate_2d(athmi_ref_grid *pG);
/* prototypes.h:112-116 (lr_states.c): nothing to allocate on the device path */
void lr_states_init(int nx1, int nx2);
void lr_states_destruct(void);
/* prototypes.h:120 (new_dt.c) */
void new_dt(athmi_ref_grid *pG);
/* prototypes.h:172-175 (set_bvals.c) */
void set_bvals_init(athmi_ref_grid *pG);
void set_bvals(athmi_ref_grid *pG, int var_flag);

/* The reference reads Gamma and CourNo from the globals of main.c
 * (globals.h:18-31) and the boundary flags from the athinput database
 * (set_bvals.c:152-230, par_geti). When the host program defines those
 * symbols they are used; otherwise this call supplies them. flux is
 * ATHMI_FLUX_ROE or ATHMI_FLUX_HLLE, variant ATHMI_EXACT or ATHMI_FAST. */
void athmi_dropin_configure(double gamma, double cour_no, int ibc_x1, int obc_x1, int ibc_x2,
                            int obc_x2, int flux, int variant);

#ifdef __cplusplus
}
#endif

#endif /* ATHENA_DROPIN_H */
