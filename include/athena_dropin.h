/*
 * athena_dropin.h -- the reference's own CPU entry points for this path, with
 * their exact names and signatures, implemented on top of athmi.h. A program
 * written against the reference (src/main.c) links libathmi.so in place of
 * __integrate_2d__.o, new_dt.o, set_bvals.o and lr_states_plm.o and runs the
 * step on the GPU. By default each call copies the Grid to the device and
 * back, so it keeps the reference's in-place semantics; resident mode (below)
 * keeps the state in HBM between the calls, and athmi_step runs the whole
 * main-loop iteration on the device.
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

/* Resident mode (off by default). The reference's GPU main loop copies the
 * Grid to the device once and keeps it there (utils_cu.c:230, main.c:221-318);
 * with resident mode on, the calls above do the same: integrate_2d, new_dt and
 * set_bvals upload the Grid only when the device does not hold its current
 * state (first call, another Grid, after athmi_dropin_mark_dirty), pass the
 * clock (dt, time, nstep) every call, and new_dt copies back dt alone. The
 * host arrays then lag the device until athmi_dropin_sync_host(pG) (call it
 * before reading them, e.g. before data_output) or until resident mode is
 * switched off with the Grid passed. A program that writes the host arrays
 * calls athmi_dropin_mark_dirty() afterwards (after syncing first, or the
 * device's newer state is lost). No reference counterpart: the reference's
 * CPU routines have no device copy to keep. */
void athmi_dropin_resident(athmi_ref_grid *pG, int on);
void athmi_dropin_sync_host(athmi_ref_grid *pG);
void athmi_dropin_mark_dirty(void);

#ifdef __cplusplus
}
#endif

#endif /* ATHENA_DROPIN_H */
