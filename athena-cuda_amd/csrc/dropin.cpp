// dropin.cpp -- the reference's CPU entry points (include/athena_dropin.h)
// over the device API. By default the Grid is copied to HBM, stepped there
// and copied back on every call, so callers keep the reference's in-place
// semantics. In resident mode (athmi_dropin_resident) the state stays in HBM
// between calls, as the reference's own GPU main loop keeps it
// (utils_cu.c:230, main.c:221-318): a call uploads only when the device does
// not hold this Grid's current state, new_dt returns dt alone, and the host
// arrays are refreshed by athmi_dropin_sync_host.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "athena_dropin.h"
#include "athmi.h"

// Process globals and the athinput lookup of a reference host program
// (globals.h:18-31, par.c:231); weak so that the library also loads alone.
extern "C" {
extern double Gamma __attribute__((weak));
extern double CourNo __attribute__((weak));
int par_geti(char* block, char* name) __attribute__((weak));
}

namespace {

struct DropinState {
  athmi_grid* g = nullptr;
  int nx1 = 0, nx2 = 0;
  double gamma = 5.0 / 3.0, cour_no = 0.4;
  int bc[4] = {ATHMI_BC_PERIODIC, ATHMI_BC_PERIODIC, ATHMI_BC_PERIODIC, ATHMI_BC_PERIODIC};
  int flux = ATHMI_FLUX_ROE, variant = ATHMI_EXACT;
  double dx1 = 0, dx2 = 0, x1_0 = 0, x2_0 = 0;
  double g_gamma = 0, g_cour = 0;  // EOS constants the handle was made with
  std::vector<double> gas, b1, b2, b3;
  bool resident = false;            // keep the state in HBM between calls
  const void* owner = nullptr;      // the Grid (its U block) the device copy belongs to
  bool dev_valid = false;           // the device holds that Grid's current state
  bool host_stale = false;          // the host arrays lag the device
};
DropinState S;

[[noreturn]] void die(const char* where) {
  // the reference reports fatal errors and exits (utils.c:111-129)
  std::fprintf(stderr, "[%s]: %s\n", where, athmi_last_error());
  std::exit(EXIT_FAILURE);
}

double eos_gamma() { return (&Gamma != nullptr) ? Gamma : S.gamma; }
double courant() { return (&CourNo != nullptr) ? CourNo : S.cour_no; }

void drop_handle() {
  if (S.g) athmi_destroy(S.g);
  S.g = nullptr;
  S.dev_valid = false;
  S.host_stale = false;
  S.owner = nullptr;
}

// The handle for this Grid's geometry and EOS; the only place that records
// the geometry it was made for, so a change of any of them makes a new one.
athmi_grid* ensure(const athmi_ref_grid* pG) {
  const bool same = S.g && S.nx1 == pG->Nx1 && S.nx2 == pG->Nx2 && S.dx1 == pG->dx1 &&
                    S.dx2 == pG->dx2 && S.x1_0 == pG->x1_0 && S.x2_0 == pG->x2_0 &&
                    S.g_gamma == eos_gamma() && S.g_cour == courant();
  if (same) return S.g;
  drop_handle();
  athmi_desc d{};
  d.nx1 = pG->Nx1;
  d.nx2 = pG->Nx2;
  d.x1min = pG->x1_0;
  d.x1max = pG->x1_0 + pG->Nx1 * pG->dx1;
  d.x2min = pG->x2_0;
  d.x2max = pG->x2_0 + pG->Nx2 * pG->dx2;
  d.gamma = eos_gamma();
  d.cour_no = courant();
  d.bc_ix1 = S.bc[0];
  d.bc_ox1 = S.bc[1];
  d.bc_ix2 = S.bc[2];
  d.bc_ox2 = S.bc[3];
  d.flux = S.flux;
  d.variant = S.variant;
  if (athmi_create(&d, 1, &S.g) != ATHMI_OK) die("integrate_init_2d");
  if (athmi_set_spacing(S.g, 0, pG->dx1, pG->dx2) != ATHMI_OK) die("integrate_init_2d");
  S.nx1 = pG->Nx1;
  S.nx2 = pG->Nx2;
  S.dx1 = pG->dx1;
  S.dx2 = pG->dx2;
  S.x1_0 = pG->x1_0;
  S.x2_0 = pG->x2_0;
  S.g_gamma = d.gamma;
  S.g_cour = d.cour_no;
  return S.g;
}

static_assert(sizeof(athmi_ref_gas) == 8 * sizeof(double), "Gas is 8 packed doubles");

// calloc_2d_array (ath_array.c:48-66) puts all rows of an array in one block:
// the Grid's arrays then have the C-ABI's layout and cross without a host copy
template <class T>
bool one_block(T* const* rows, int sy, int sx) {
  for (int j = 1; j < sy; j++)
    if (rows[j] != rows[0] + (size_t)j * sx) return false;
  return true;
}

bool grid_is_one_block(const athmi_ref_grid* pG) {
  const int sx = pG->Nx1 + 8, sy = pG->Nx2 + 8;
  return one_block(pG->U, sy, sx) && one_block(pG->B1i, sy, sx) &&
         one_block(pG->B2i, sy, sx) && one_block(pG->B3i, sy, sx);
}

// Make the device hold pG's state and clock. The arrays cross only when the
// device does not already hold this Grid's current state (always, outside
// resident mode); the clock (three scalars the caller advances itself,
// main.c:266-271) is written every time.
void to_device(const athmi_ref_grid* pG) {
  athmi_grid* g = ensure(pG);
  if (S.resident && S.dev_valid && S.owner == (const void*)pG->U[0]) {
    if (athmi_set_clock(g, 0, pG->dt, pG->time, pG->nstep) != ATHMI_OK) die("athmi_set_clock");
    return;
  }
  const int sx = pG->Nx1 + 8, sy = pG->Nx2 + 8;
  S.owner = (const void*)pG->U[0];
  S.dev_valid = true;
  S.host_stale = false;
  if (grid_is_one_block(pG)) {
    if (athmi_upload(g, 0, &pG->U[0][0].d, pG->B1i[0], pG->B2i[0], pG->B3i[0]) != ATHMI_OK)
      die("athmi_upload");
    if (athmi_set_clock(g, 0, pG->dt, pG->time, pG->nstep) != ATHMI_OK) die("athmi_set_clock");
    return;
  }
  S.gas.resize((size_t)8 * sx * sy);
  S.b1.resize((size_t)sx * sy);
  S.b2.resize((size_t)sx * sy);
  S.b3.resize((size_t)sx * sy);
  for (int j = 0; j < sy; j++)
    for (int i = 0; i < sx; i++) {
      const size_t h = (size_t)j * sx + i;
      const athmi_ref_gas& u = pG->U[j][i];
      const double v[8] = {u.d, u.M1, u.M2, u.M3, u.E, u.B1c, u.B2c, u.B3c};
      for (int q = 0; q < 8; q++) S.gas[8 * h + q] = v[q];
      S.b1[h] = pG->B1i[j][i];
      S.b2[h] = pG->B2i[j][i];
      S.b3[h] = pG->B3i[j][i];
    }
  if (athmi_upload(g, 0, S.gas.data(), S.b1.data(), S.b2.data(), S.b3.data()) != ATHMI_OK)
    die("athmi_upload");
  if (athmi_set_clock(g, 0, pG->dt, pG->time, pG->nstep) != ATHMI_OK) die("athmi_set_clock");
}

void download_now(athmi_ref_grid* pG) {
  const int sx = pG->Nx1 + 8, sy = pG->Nx2 + 8;
  S.host_stale = false;
  if (grid_is_one_block(pG)) {
    if (athmi_download(S.g, 0, &pG->U[0][0].d, pG->B1i[0], pG->B2i[0], pG->B3i[0]) != ATHMI_OK)
      die("athmi_download");
    return;
  }
  if (athmi_download(S.g, 0, S.gas.data(), S.b1.data(), S.b2.data(), S.b3.data()) != ATHMI_OK)
    die("athmi_download");
  for (int j = 0; j < sy; j++)
    for (int i = 0; i < sx; i++) {
      const size_t h = (size_t)j * sx + i;
      athmi_ref_gas& u = pG->U[j][i];
      const double* v = &S.gas[8 * h];
      u.d = v[0]; u.M1 = v[1]; u.M2 = v[2]; u.M3 = v[3];
      u.E = v[4]; u.B1c = v[5]; u.B2c = v[6]; u.B3c = v[7];
      pG->B1i[j][i] = S.b1[h];
      pG->B2i[j][i] = S.b2[h];
      pG->B3i[j][i] = S.b3[h];
    }
}

// After a device call that changed the state: copy it back now, or (resident
// mode) leave the host arrays stale until athmi_dropin_sync_host.
void from_device(athmi_ref_grid* pG) {
  if (S.resident) {
    S.host_stale = true;
    return;
  }
  download_now(pG);
}

}  // namespace

extern "C" {

void athmi_dropin_configure(double gamma, double cour_no, int ibc_x1, int obc_x1, int ibc_x2,
                            int obc_x2, int flux, int variant) {
  S.gamma = gamma;
  S.cour_no = cour_no;
  S.bc[0] = ibc_x1;
  S.bc[1] = obc_x1;
  S.bc[2] = ibc_x2;
  S.bc[3] = obc_x2;
  S.flux = flux;
  S.variant = variant;
  drop_handle();
}

void athmi_dropin_resident(athmi_ref_grid* pG, int on) {
  if (!on && S.resident && S.host_stale && pG) download_now(pG);
  S.resident = on != 0;
  if (!S.resident) S.dev_valid = false;
}

void athmi_dropin_sync_host(athmi_ref_grid* pG) {
  if (pG && S.g && S.host_stale && S.owner == (const void*)pG->U[0]) download_now(pG);
}

void athmi_dropin_mark_dirty(void) { S.dev_valid = false; }

// The device handle is made on the first call that passes a Grid (its
// geometry and EOS decide it), so there is nothing to size here.
void integrate_init_2d(int Nx1, int Nx2) {
  (void)Nx1;
  (void)Nx2;
}

void integrate_destruct_2d(void) { drop_handle(); }

void lr_states_init(int, int) {}
void lr_states_destruct(void) {}

void integrate_2d(athmi_ref_grid* pG) {
  to_device(pG);
  if (athmi_integrate(S.g) != ATHMI_OK) die("integrate_2d");
  from_device(pG);
}

void new_dt(athmi_ref_grid* pG) {
  to_device(pG);
  if (athmi_new_dt(S.g) != ATHMI_OK) die("new_dt");
  double dt;
  if (athmi_get_clock(S.g, 0, &dt, nullptr, nullptr) != ATHMI_OK) die("new_dt");
  pG->dt = dt;
}

void set_bvals_init(athmi_ref_grid* pG) {
  (void)pG;
  if (par_geti != nullptr) {
    S.bc[0] = par_geti((char*)"grid", (char*)"ibc_x1");
    S.bc[1] = par_geti((char*)"grid", (char*)"obc_x1");
    S.bc[2] = par_geti((char*)"grid", (char*)"ibc_x2");
    S.bc[3] = par_geti((char*)"grid", (char*)"obc_x2");
    drop_handle();
  }
}

void set_bvals(athmi_ref_grid* pG, int var_flag) {
  if (var_flag == 1) return;  // every boundary function returns for var_flag 1
  to_device(pG);
  if (athmi_set_bvals(S.g) != ATHMI_OK) die("set_bvals");
  from_device(pG);
}

}  // extern "C"
