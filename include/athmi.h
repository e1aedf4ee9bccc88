/*
 * athmi.h -- C-ABI of the MI355X-native 2-D CTU + constrained-transport MHD
 * step (the Athena 3.1 path of awasiljew/athena-cuda).
 *
 * One handle holds one or more independent grid instances resident in HBM
 * (a plain grid is a batch of one). Every function returns ATHMI_OK (0) or a
 * negative error code; athmi_last_error() describes the last failure of the
 * calling thread. No HIP, CUDA or torch types appear: streams are passed as
 * void*. Each entry point cites the reference interface it replaces.
 *
 * Host arrays use the reference layouts: the conserved state is the Gas
 * array of athena.h:43-52 (8 doubles per cell: d, M1, M2, M3, E, B1c, B2c,
 * B3c), row-major [Nx2+8][Nx1+8] including the 4 ghost layers; the
 * interface fields B1i, B2i, B3i are [Nx2+8][Nx1+8] doubles
 * (athena.h:106-109, init_grid.c:170-210). Index (j,i) = cell (i,j).
 */
#ifndef ATHMI_H
#define ATHMI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATHMI_OK 0
#define ATHMI_EINVAL (-1)  /* bad argument */
#define ATHMI_EHIP (-2)    /* HIP runtime error */
#define ATHMI_ENOMEM (-3)  /* device or host allocation failed */
#define ATHMI_ESTATE (-4)  /* call not valid in the handle's current state */

/* boundary flags of the athinput <grid> block (set_bvals.c:26-29) */
#define ATHMI_BC_REFLECT 1
#define ATHMI_BC_OUTFLOW 2
#define ATHMI_BC_PERIODIC 4

#define ATHMI_FLUX_ROE 0  /* flux_roe.c:54 with its HLLE fallback */
#define ATHMI_FLUX_HLLE 1 /* flux_hlle.c:60 */

#define ATHMI_EXACT 0 /* one rounding per operation (no FMA contraction) */
#define ATHMI_FAST 1  /* FMA-contracted arithmetic; division and square root from
                         the reciprocal estimates refined by Newton steps
                         (about an ulp) */

typedef struct athmi_grid athmi_grid;

/* Geometry, EOS and options of one instance; mirrors init_grid.c:25-118 plus
 * the <time>/<problem> constants read by main.c:64-94. */
typedef struct athmi_desc {
  int nx1, nx2;                 /* interior zones (Nx1, Nx2 >= 2) */
  double x1min, x1max, x2min, x2max;
  double gamma;                 /* Gamma; Gamma_1, Gamma_2 are derived */
  double cour_no;               /* CourNo */
  int bc_ix1, bc_ox1, bc_ix2, bc_ox2; /* ATHMI_BC_* */
  int flux;                     /* ATHMI_FLUX_* */
  int variant;                  /* ATHMI_EXACT or ATHMI_FAST (same for a whole batch) */
} athmi_desc;

/* Allocate n instances (device state, scratch and clocks).
 * Replaces init_grid_gpu + integrate_init_2d_cu + lr_states_init_cu + dt_init
 * + set_bvals_init_cu (utils_cu.c:132, integrate_2d_cuda.cu:218,333,
 * new_dt_cuda.cu:241, set_bvals_cuda.cu:591). */
int athmi_create(const athmi_desc *desc, int n, athmi_grid **out);
/* Replaces integrate_destruct_2d_cu + dt_destruct + the cudaFree calls of
 * main.c:321-345. */
int athmi_destroy(athmi_grid *g);
int athmi_count(const athmi_grid *g, int *n);
/* Use exactly these cell sizes (a caller holding a Grid built elsewhere,
 * whose dx may not round-trip through (max - min)/N) */
int athmi_set_spacing(athmi_grid *g, int inst, double dx1, double dx2);

/* Host <-> device copies of one instance's full padded state (AoS <-> SoA).
 * Replace copy_to_gpu_mem / copy_to_host_mem (utils_cu.c:230, 293) without
 * their 4-cell x shift. The arrays cross PCIe unchanged and the layout
 * conversion runs on the device, staged in the instance's idle ping-pong
 * buffer (the next step rewrites it). Both synchronise. */
int athmi_upload(athmi_grid *g, int inst, const double *gas, const double *b1i,
                 const double *b2i, const double *b3i);
int athmi_download(athmi_grid *g, int inst, double *gas, double *b1i, double *b2i,
                   double *b3i);

/* Grid.dt / Grid.time / Grid.nstep (athena.h:113-114) of one instance */
int athmi_set_clock(athmi_grid *g, int inst, double dt, double time, int nstep);
int athmi_get_clock(athmi_grid *g, int inst, double *dt, double *time, int *nstep);
/* Failure word of one instance: the step checks every state it produces and
 * raises ATHMI_FAIL_* bits. Once a bit of the instance's stop set is raised,
 * the main-loop step (athmi_step, athmi_run) no longer runs on that instance;
 * the others go on. *nstep gets the step number of the first state that
 * raised a bit (-1: none). The reference has no such check: it floors the
 * pressure (convert_var.c:49, new_dt.c:54-55) and steps on, so the default
 * stop set is NONFINITE | DENSITY (the states from which no further step
 * means anything) and a pressure-floor hit is only reported -- the run goes on
 * exactly as the reference's. athmi_upload of a new state clears the word.
 * Either pointer may be NULL. */
#define ATHMI_FAIL_NONFINITE 1 /* NaN or infinity in d, M1..M3, E or cell-centred B */
#define ATHMI_FAIL_DENSITY 2   /* d <= 0 */
#define ATHMI_FAIL_PRESSURE 4  /* Gamma_1 (E - KE - ME) < TINY_NUMBER (the floor engages) */
int athmi_get_failure(athmi_grid *g, int inst, int *flags, int *nstep);
/* The stop set of one instance (inst = -1: all), any OR of ATHMI_FAIL_*;
 * 0 never stops. No reference counterpart. */
int athmi_set_fail_stop(athmi_grid *g, int inst, int mask);
/* main-loop limits (main.c:217-218: tlim, nlim; nlim < 0 = none); inst = -1: all */
int athmi_set_limits(athmi_grid *g, int inst, double tlim, int nlim);
/* Output cadence (main.c:226 calls data_output, output.c:362-376, at the top
 * of every iteration): the main-loop step of an instance does not run once
 * its time >= t_next, so that the caller can write the outputs due
 * (athmi_output.h) before the step; HUGE_VAL (the default) never waits.
 * inst = -1: all instances. */
int athmi_set_pause(athmi_grid *g, int inst, double t_next);

/* set_bvals(Grid*, 0) (set_bvals.c:114) on every instance */
int athmi_set_bvals(athmi_grid *g);
/* new_dt(Grid*) (new_dt.c:25) on every instance; result in the clock's dt */
int athmi_new_dt(athmi_grid *g);
/* integrate_2d(Grid*) (__integrate_2d__.c:100) on every instance with its
 * clock's dt; the clock's time and nstep are left unchanged */
int athmi_integrate(athmi_grid *g);
/* One iteration of the main loop (main.c:221-318) on every instance that is
 * still running: clamp dt to tlim, integrate, nstep++, time += dt, new_dt,
 * set_bvals. Everything stays on the device; replayed from a HIP graph. */
int athmi_step(athmi_grid *g);
/* Repeat athmi_step until no instance is running or max_steps iterations
 * (max_steps < 0: no cap); *steps = iterations issued. Synchronises. */
int athmi_run(athmi_grid *g, int max_steps, int *steps);
/* max |div B| over the interior of one instance */
int athmi_divb_max(athmi_grid *g, int inst, double *out);

/* Execution control */
int athmi_set_stream(athmi_grid *g, void *hip_stream); /* NULL: default stream */
int athmi_sync(athmi_grid *g);
/* Launch shape: faces marched per x2-predictor thread, persistent
 * workgroups per CU (0 = one tile per workgroup). Picked by the autotuner;
 * the default marches about sqrt(C)/50 faces (1..16), C the x2 face columns
 * of 64 lanes (1 at 256^2, 3 at 1024^2, 5 at 2048^2, 10 at 4096^2). */
int athmi_set_shape(athmi_grid *g, int x2_seg, int persist);
/* Tile order: 1 = each of the 8 XCDs gets one contiguous run of tiles (its L2
 * serves the neighbour rows), 0 = launch order. Results are identical. */
int athmi_set_tile_order(athmi_grid *g, int xcd_remap);
/* Rows of a corrector tile (64 x rows faces: 1, 2 or 4; default 2). Results
 * are identical. */
int athmi_set_corr_rows(athmi_grid *g, int rows);
/* Form of the x2 predictor: 0 = each lane marches its column with the PLM
 * stencil in registers (one wave per SIMD), 1 = the stencil and pending face
 * states in LDS slots (two waves per SIMD). Picked by the autotuner; results
 * are identical. */
int athmi_set_x2_kind(athmi_grid *g, int kind);
int athmi_get_x2_kind(const athmi_grid *g, int *kind);
/* Fusion depth of the step: 0 = the corner emf of Steps 4 and 9
 * (integrate_emf3_corner, __integrate_2d__.c:632-702) each in its own kernel,
 * written to a plane the next kernel reads; 1 = the Step 9 corners computed
 * inside the final update's tiles; 2 = also the Step 4 corners inside the x1
 * corrector's tiles. Picked by the autotuner; results are identical. */
int athmi_set_fusion(athmi_grid *g, int depth);
int athmi_get_fusion(const athmi_grid *g, int *depth);
/* Tile order of the corrector and final kernels: 0 = row by row, s > 0 =
 * strips of s tile columns (64 s cells wide), each walked row by row, so that
 * the tile dispatched just before a tile is mostly the one above it and the
 * halo rows they share are still in the XCD's L2. Picked by the autotuner;
 * results are identical. */
int athmi_set_strip(athmi_grid *g, int tile_cols);
int athmi_get_strip(const athmi_grid *g, int *tile_cols);
/* Rows of a final-update tile (64 x rows cells: 4, 8 or 16; default 8): the
 * taller the tile, the smaller the share of its halo rows read twice.
 * Picked by the autotuner; results are identical. */
int athmi_set_final_rows(athmi_grid *g, int rows);
int athmi_get_final_rows(const athmi_grid *g, int *rows);
/* The current launch shape (any pointer may be NULL) */
int athmi_get_shape(const athmi_grid *g, int *x2_seg, int *persist, int *xcd_remap,
                    int *corr_rows);
int athmi_use_graph(athmi_grid *g, int on);
/* Autotuner: time every launch shape (x2 predictor form x x2 marching depth
 * 2/4/8/16 x persistent workgroups per CU 0..32 x tile order, then the x2
 * marching depth 1..32, the corrector's tile height 4/2/1, the corrector and
 * final kernels' strip
 * width 0/2/4/8, the final tile height 4/8/16 and the fusion depth at the
 * best of those)
 * by the better of two runs of `steps` main-loop steps from the current state
 * through the step graph, restore the state and clocks; the six fastest
 * shapes then run three more trials each, in turn, and the lowest total wins
 * (returned with its mean ms per step; athmi_get_shape and the getters give
 * all of it). No reference counterpart (the reference launches fixed 32-thread
 * blocks, integrate_2d_cuda.cu:93-193). */
int athmi_autotune(athmi_grid *g, int steps, int *x2_seg, int *persist, int *xcd_remap,
                   double *ms_per_step);
/* Time nsteps main-loop steps kernel by kernel (HIP events on the handle's
 * stream between consecutive launches, no graph); avg_ms[k] receives the mean
 * duration of kernel k of the step (nslots >= 10), named by
 * athmi_step_kernel_name(k). No reference counterpart: measurement hook. */
int athmi_profile_steps(athmi_grid *g, int nsteps, double *avg_ms, int nslots);
const char *athmi_step_kernel_name(int k);
/* Copy nbytes between two fresh device buffers reps times with a plain
 * streaming kernel (k_copy_calibration): the reference traffic that scales
 * rocprofv3's FETCH_SIZE / WRITE_SIZE. No reference counterpart. */
int athmi_copy_calibration(size_t nbytes, int reps);

/* Device pointers of the current state of one instance: 11 planes
 * (d, M1, M2, M3, E, B1c, B2c, B3c, B1i, B2i, B3i), each [Nx2+8][pitch] */
int athmi_device_state(athmi_grid *g, int inst, double **planes, int *pitch);

/* Problem generators (host libm, as the reference), applied to one instance's
 * interior; follow with athmi_set_bvals and athmi_new_dt as main.c:121-151.
 * blast.c:25-72 (cx, cy: centre of the over-pressured region, 0,0 in the
 * reference); field_loop.c:40-137; cpaw2d.c:56-185 (returns the wavelength;
 * keeps the initial state for athmi_cpaw2d_l1). */
int athmi_pgen_blast(athmi_grid *g, int inst, double radius, double pamb, double prat,
                     double b0, double angle_deg, double cx, double cy);
int athmi_pgen_field_loop(athmi_grid *g, int inst, double rad, double amp, double vflow,
                          int iprob);
int athmi_pgen_cpaw2d(athmi_grid *g, int inst, double angle, double b_par, double b_perp,
                      double pres, double v_par, double *lambda);
/* The same generators on the device, for instances [first, first + n) in one
 * launch over their cells (the batch's initial states never cross PCIe).
 * params[m] holds the problem and its <problem> parameters in the order of the
 * host functions above: blast (radius, pamb, prat, b0, angle, cx, cy),
 * field_loop (rad, amp, vflow, iprob), cpaw2d (angle, b_par, b_perp, pres,
 * v_par); lambda[m] (may be NULL) receives cpaw2d's wavelength. Per-grid
 * constants use the host libm: blast and field_loop states are identical to
 * the host generators'; cpaw2d's per-cell sin/cos use the device libm (within
 * an ulp). The host generators stay the parity mode. */
enum { ATHMI_PGEN_BLAST = 0, ATHMI_PGEN_FIELD_LOOP = 1, ATHMI_PGEN_CPAW2D = 2 };
typedef struct athmi_pgen_params {
  int problem;
  double p[7];
} athmi_pgen_params;
int athmi_pgen_device(athmi_grid *g, int first, int n, const athmi_pgen_params *params,
                      double *lambda);

/* cpaw2d.c:222-302: out[0] = RMS error, out[1..8] = L1 error per variable in
 * Gas order (d, M1, M2, M3, E, B1c, B2c, B3c), each divided by Nx1*Nx2 */
int athmi_cpaw2d_l1(athmi_grid *g, int inst, double *out);

/* integrate_2d of one grid on caller-owned device planes, allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP stream capture, tight serving loops, harnesses that
 * forbid hipMalloc). state_in and state_out each hold the 11 planes (d, M1,
 * M2, M3, E, B1c, B2c, B3c, B1i, B2i, B3i) of [Nx2+8][pitch] doubles,
 * plane_stride doubles apart, ghost zones included; state_in is read,
 * state_out receives the state athmi_integrate would produce with this dt
 * (interior updated, ghost zones carried unchanged). The work is queued on
 * `hip_stream` (NULL: the null stream) and not synchronised. The descriptor,
 * clock and the 59 scratch planes live in one static device array of
 * ATHMI_STATIC_SCRATCH_MB MiB per GPU, held until the process ends; a grid
 * needs 4 KiB + 59 * plane_stride * 8 bytes of it (plane_stride up to about
 * 284k doubles: e.g. 400x600 with pitch 416) and larger ones are rejected
 * with ATHMI_EINVAL. Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other threads) race. */
#define ATHMI_STATIC_SCRATCH_MB 128
int athmi_integrate_static(const athmi_desc *desc, double dt, const double *state_in,
                           double *state_out, int pitch, size_t plane_stride, void *hip_stream);

const char *athmi_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ATHMI_H */
