/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A thin C driver around the reference's own CPU sources, compiled from
 * /root/reference/src by oracle/Makefile into oracle/_ref/libathref_{roe,hlle}.so.
 * It reproduces the ONLY_HOST main loop of src/main.c:221-318 (the reference
 * main.c itself needs the CUDA runtime and cannot be linked here) and exposes
 * single reference routines so that tests can feed them arbitrary states.
 *
 * Only tests/, tests/golden/ scripts and bench.py's cpu_baseline leg load it.
 */
#define MAIN_C
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "defs.h"
#include "athena.h"
#include "globals.h"
#include "prototypes.h"

/* problem generators, renamed per translation unit by oracle/Makefile */
void problem_blast(Grid *pG);
void problem_field_loop(Grid *pG);
void problem_cpaw2d(Grid *pG);
void userwork_after_loop_cpaw2d(Grid *pG);
Gasfun_t gue_blast(const char *expr);
Gasfun_t gue_field_loop(const char *expr);
Gasfun_t gue_cpaw2d(const char *expr);

static Grid G;
static int have_grid = 0;
static int par_is_open = 0;
static int cur_problem = -1;
static char outname[256] = "ref";

/* main.c:64-94: parameters, Courant number and EOS constants */
int ref_open(const char *athinput, int nover, const char **over)
{
  int i;
  char **argv = (char **)calloc((size_t)nover + 1, sizeof(char *));
  if (par_is_open) par_close();
  par_open((char *)athinput);
  par_is_open = 1;
  argv[0] = "ref_harness";
  for (i = 0; i < nover; i++) argv[i + 1] = strdup(over[i]);
  par_cmdline(nover + 1, argv);
  for (i = 0; i < nover; i++) free(argv[i + 1]);
  free(argv);
  CourNo = par_getd("time", "cour_no");
  Gamma = par_getd("problem", "gamma");
  Gamma_1 = Gamma - 1.0;
  Gamma_2 = Gamma - 2.0;
  return 0;
}

/* EOS constants for single-routine probes that need no athinput */
void ref_set_gamma(double gamma)
{
  Gamma = gamma;
  Gamma_1 = Gamma - 1.0;
  Gamma_2 = Gamma - 2.0;
}

/* main.c:96-207 restricted to the ONLY_HOST path */
int ref_setup(int problem)
{
  memset(&G, 0, sizeof(G));
  G.outfilename = outname;
  cur_problem = problem;
  init_grid(&G);
  if (problem == 0) problem_blast(&G);
  else if (problem == 1) problem_field_loop(&G);
  else if (problem == 2) problem_cpaw2d(&G);
  else return 1;
  set_bvals_init(&G);
  set_bvals(&G, 0);
  new_dt(&G);
  lr_states_init(G.Nx1, G.Nx2);
  integrate_init_2d(G.Nx1, G.Nx2);
  have_grid = 1;
  return 0;
}

void ref_info(int *iv, double *dv)
{
  iv[0] = G.Nx1; iv[1] = G.Nx2; iv[2] = G.is; iv[3] = G.ie;
  iv[4] = G.js; iv[5] = G.je; iv[6] = G.nstep; iv[7] = G.idisp; iv[8] = G.jdisp;
  dv[0] = G.x1_0; dv[1] = G.x2_0; dv[2] = G.dx1; dv[3] = G.dx2;
  dv[4] = G.dt; dv[5] = G.time;
}

/* one iteration of main.c:221-318 (ONLY_HOST): clamp, integrate, advance, dt, BCs */
int ref_step(double tlim)
{
  if (!have_grid) return 1;
  if ((tlim - G.time) < G.dt) G.dt = (tlim - G.time);
  integrate_2d(&G);
  G.nstep++;
  G.time += G.dt;
  new_dt(&G);
  set_bvals(&G, 0);
  return 0;
}

/* run the loop to tlim or nlim steps; returns the number of steps taken */
int ref_run(double tlim, int nlim)
{
  int n = 0;
  while (G.time < tlim && (nlim < 0 || n < nlim)) {
    ref_step(tlim);
    n++;
  }
  return n;
}

/* The reference's output layer (output.c, dump_binary.c, output_ppm.c,
 * dump_tab.c, ath_files.c) on the harness grid. Files go to the current
 * directory as <outfilename>[.num][.id].<ext>. */
Gasfun_t get_usr_expr(const char *expr)  /* the linked problem's (output.c:233) */
{
  if (cur_problem == 0) return gue_blast(expr);
  if (cur_problem == 1) return gue_field_loop(expr);
  if (cur_problem == 2) return gue_cpaw2d(expr);
  return NULL;
}

void ref_set_outfilename(const char *name)  /* main.c:90 (job/problem_id) */
{
  strncpy(outname, name, sizeof(outname) - 1);
  outname[sizeof(outname) - 1] = 0;
}

/* getRGB (output.c:862-880); name must be one of the palettes of palette.h */
float *getRGB(char *name);
void ref_get_palette(const char *name, float *out768)
{
  const float *rgb = getRGB((char *)name);
  memcpy(out768, rgb, 768 * sizeof(float));
}

void ref_init_output(void) { init_output(&G); }
void ref_data_output(int flag) { data_output(&G, NULL, flag); }
void ref_output_destruct(void) { data_output_destruct(); }

/* main.c:221-318 (ONLY_HOST) with data_output at the top of each iteration
 * (main.c:226); returns the number of steps taken */
int ref_run_with_output(double tlim, int nlim)
{
  int n = 0;
  if (!have_grid) return -1;
  while (G.time < tlim && (nlim < 0 || G.nstep < nlim)) {
    data_output(&G, NULL, 0);
    ref_step(tlim);
    n++;
  }
  return n;
}

/* single routines of the step, for staged comparisons */
void ref_integrate(void) { integrate_2d(&G); }
void ref_new_dt(void) { new_dt(&G); }
void ref_set_bvals(void) { set_bvals(&G, 0); }

void ref_set_time(double dt, double time, int nstep)
{
  G.dt = dt; G.time = time; G.nstep = nstep;
}

/* U is AoS Gas[sy][sx] (8 doubles per cell); B*i are [sy][sx] */
void ref_get(double *U, double *B1, double *B2, double *B3)
{
  int sx = G.Nx1 + 2 * nghost, sy = G.Nx2 + 2 * nghost;
  memcpy(U, G.U[0], sizeof(Gas) * (size_t)sx * sy);
  memcpy(B1, G.B1i[0], sizeof(Real) * (size_t)sx * sy);
  memcpy(B2, G.B2i[0], sizeof(Real) * (size_t)sx * sy);
  memcpy(B3, G.B3i[0], sizeof(Real) * (size_t)sx * sy);
}

void ref_put(const double *U, const double *B1, const double *B2, const double *B3)
{
  int sx = G.Nx1 + 2 * nghost, sy = G.Nx2 + 2 * nghost;
  memcpy(G.U[0], U, sizeof(Gas) * (size_t)sx * sy);
  memcpy(G.B1i[0], B1, sizeof(Real) * (size_t)sx * sy);
  memcpy(G.B2i[0], B2, sizeof(Real) * (size_t)sx * sy);
  memcpy(G.B3i[0], B3, sizeof(Real) * (size_t)sx * sy);
}

/* cpaw2d.c:222-302 writes cpaw2d-errors.dat into the working directory */
int ref_cpaw2d_errors(const char *dir)
{
  char cwd[4096];
  if (getcwd(cwd, sizeof(cwd)) == NULL) return 1;
  if (chdir(dir) != 0) return 2;
  userwork_after_loop_cpaw2d(&G);
  if (chdir(cwd) != 0) return 3;
  return 0;
}

/* ---- single reference routines ---------------------------------------- */

void ref_esys_prim(const double *in, double *ev, double *rem, double *lem)
{
  double r[7][7], l[7][7];
  memset(r, 0, sizeof(r));
  memset(l, 0, sizeof(l));
  esys_prim_adb_mhd(in[0], in[1], in[2], in[3], in[4], in[5], ev, r, l);
  memcpy(rem, r, sizeof(r));
  memcpy(lem, l, sizeof(l));
}

void ref_esys_roe(const double *in, double *ev, double *rem, double *lem, int eig_only)
{
  double r[7][7], l[7][7];
  memset(r, 0, sizeof(r));
  memset(l, 0, sizeof(l));
  if (eig_only)
    esys_roe_adb_mhd(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7],
                     in[8], in[9], ev, NULL, NULL);
  else
    esys_roe_adb_mhd(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7],
                     in[8], in[9], ev, r, l);
  memcpy(rem, r, sizeof(r));
  memcpy(lem, l, sizeof(l));
}

void ref_flux_roe(const double *ul, const double *ur, const double *wl,
                  const double *wr, double bxi, double *f)
{
  Cons1D Ul, Ur, F;
  Prim1D Wl, Wr;
  memcpy(&Ul, ul, sizeof(Ul)); memcpy(&Ur, ur, sizeof(Ur));
  memcpy(&Wl, wl, sizeof(Wl)); memcpy(&Wr, wr, sizeof(Wr));
  flux_roe(Ul, Ur, Wl, Wr, bxi, &F);
  memcpy(f, &F, sizeof(F));
}

void ref_flux_hlle(const double *ul, const double *ur, const double *wl,
                   const double *wr, double bxi, double *f)
{
  Cons1D Ul, Ur, F;
  Prim1D Wl, Wr;
  memcpy(&Ul, ul, sizeof(Ul)); memcpy(&Ur, ur, sizeof(Ur));
  memcpy(&Wl, wl, sizeof(Wl)); memcpy(&Wr, wr, sizeof(Wr));
  flux_hlle(Ul, Ur, Wl, Wr, bxi, &F);
  memcpy(f, &F, sizeof(F));
}

void ref_cons_to_prim(const double *u, double bx, double *w)
{
  Cons1D U; Prim1D W;
  memcpy(&U, u, sizeof(U));
  Cons1D_to_Prim1D(&U, &W, &bx);
  memcpy(w, &W, sizeof(W));
}

void ref_prim_to_cons(const double *w, double bx, double *u)
{
  Cons1D U; Prim1D W;
  memcpy(&W, w, sizeof(W));
  Prim1D_to_Cons1D(&U, &W, &bx);
  memcpy(u, &U, sizeof(U));
}

/* lr_states over a slice of n cells; W is Prim1D[n], Wl/Wr are Prim1D[n] */
void ref_lr_states(int n, const double *W, const double *Bxc, double dt,
                   double dtodx, int il, int iu, double *Wl, double *Wr)
{
  /* lr_states_init (re)allocates the slice pointer table; the old one leaks,
   * which is harmless in a probe process */
  lr_states_init(n, n);
  lr_states((const Prim1D *)W, Bxc, dt, dtodx, il, iu, (Prim1D *)Wl, (Prim1D *)Wr);
}

void ref_cc_pos(int i, int j, double *x1, double *x2)
{
  cc_pos(&G, i, j, x1, x2);
}
