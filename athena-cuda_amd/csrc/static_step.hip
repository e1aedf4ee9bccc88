// static_step.hip -- integrate_2d on caller-owned device planes with every
// working buffer (instance descriptor, clock, the 59 scratch planes) in one
// file-scope device array: the entry point allocates nothing, for callers
// that cannot call hipMalloc (code under stream capture, tight serving loops).
//
// Each GPU the library runs on holds its own copy of `scratch` for the life
// of the process. The host never passes the array's name to a kernel: it
// resolves the array's device address once per device (hipGetSymbolAddress,
// with that device current) and passes that address. Every call on a device
// shares its copy; calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "athmi.h"
#include "runtime.hpp"
#include "step_kernels.hpp"

namespace {

constexpr size_t kStaticBytes = (size_t)ATHMI_STATIC_SCRATCH_MB << 20;
constexpr size_t kHeadBytes = 4096;  // Inst, then Clock at +1024
static_assert(sizeof(athmi::Inst) <= 1024, "Inst outgrew its slot");
static_assert(1024 + sizeof(athmi::Clock) <= kHeadBytes, "Clock outgrew its slot");

__device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices];  // device address of `scratch` per device
int g_ncu[kMaxDevices];

// Descriptor and clock of the one instance, written by a kernel so that the
// call holds no host-to-device copy (a captured copy would read the host
// value at replay time)
__global__ void k_static_head(athmi::Inst inst, athmi::Clock* clk_dst, athmi::Inst* inst_dst,
                              double dt, double cour_no) {
  if (threadIdx.x == 0) *inst_dst = inst;
  // the whole clock starts at zero, then the head fields a single
  // integration reads: dt, the ping-pong index (0: input), no failure
  unsigned long long* w = reinterpret_cast<unsigned long long*>(clk_dst);
  for (size_t e = threadIdx.x; e < sizeof(athmi::Clock) / 8; e += blockDim.x) w[e] = 0ull;
  __syncthreads();
  if (threadIdx.x == 0) {
    clk_dst->dt = dt;
    clk_dst->cour_no = cour_no;
    clk_dst->tlim = HUGE_VAL;
    clk_dst->tpause = HUGE_VAL;
    clk_dst->nlim = -1;
    clk_dst->fail_nstep = -1;
    clk_dst->active = 1;
    clk_dst->cur = 0;
  }
}

// zero n doubles (the scratch planes this call uses keep no earlier call's data)
__global__ void __launch_bounds__(256) k_static_clear(double* p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    p[e] = 0.0;
}

int resolve(int dev, unsigned char** ptr, int* ncu) {
  if (dev < 0 || dev >= kMaxDevices)
    return athmi_rt::fail(ATHMI_EINVAL, "athmi_integrate_static: device index too large");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return athmi_rt::fail(ATHMI_EHIP, std::string("athmi_integrate_static: "
                                                    "hipGetSymbolAddress: ") +
                                            hipGetErrorString(e));
    int n = 0;
    e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess)
      return athmi_rt::fail(ATHMI_EHIP, std::string("athmi_integrate_static: "
                                                    "hipDeviceGetAttribute: ") +
                                            hipGetErrorString(e));
    g_ncu[dev] = n;
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *ptr = g_ptr[dev];
  *ncu = g_ncu[dev];
  return ATHMI_OK;
}

}  // namespace

extern "C" {

int athmi_integrate_static(const athmi_desc* desc, double dt, const double* state_in,
                           double* state_out, int pitch, size_t plane_stride, void* stream) {
  using namespace athmi;
  using athmi_rt::fail;
  if (!desc || !state_in || !state_out)
    return fail(ATHMI_EINVAL, "athmi_integrate_static: null argument");
  const athmi_desc& d = *desc;
  if (d.nx1 < 2 || d.nx2 < 2)
    return fail(ATHMI_EINVAL, "athmi_integrate_static: the 2-D integrator needs Nx1, Nx2 > 1");
  if (!(d.x1max > d.x1min) || !(d.x2max > d.x2min))
    return fail(ATHMI_EINVAL, "athmi_integrate_static: empty domain");
  const int bcs[4] = {d.bc_ix1, d.bc_ox1, d.bc_ix2, d.bc_ox2};
  for (int b : bcs)
    if (b != ATHMI_BC_REFLECT && b != ATHMI_BC_OUTFLOW && b != ATHMI_BC_PERIODIC)
      return fail(ATHMI_EINVAL, "athmi_integrate_static: boundary flag must be 1, 2 or 4");
  if (d.flux != ATHMI_FLUX_ROE && d.flux != ATHMI_FLUX_HLLE)
    return fail(ATHMI_EINVAL, "athmi_integrate_static: unknown flux");
  if (d.variant != ATHMI_EXACT && d.variant != ATHMI_FAST)
    return fail(ATHMI_EINVAL, "athmi_integrate_static: unknown variant");
  const int sx = d.nx1 + 8, sy = d.nx2 + 8;
  if (pitch < sx || plane_stride < (size_t)pitch * sy)
    return fail(ATHMI_EINVAL, "athmi_integrate_static: pitch < Nx1 + 8 or plane stride < "
                              "pitch * (Nx2 + 8)");
  // plane-element offsets are 32-bit in the kernels
  if ((size_t)pitch * sy >= ((size_t)1 << 29))
    return fail(ATHMI_EINVAL, "athmi_integrate_static: plane too large");
  const size_t need = kHeadBytes + (size_t)NSCRATCH * plane_stride * sizeof(double);
  if (need > kStaticBytes)
    return fail(ATHMI_EINVAL,
                "athmi_integrate_static: " + std::to_string(NSCRATCH) + " scratch planes of " +
                    std::to_string(plane_stride) + " doubles need " + std::to_string(need) +
                    " bytes; the static workspace holds " + std::to_string(kStaticBytes) +
                    " (ATHMI_STATIC_SCRATCH_MB)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  int ncu = 0;
  int rc = resolve(dev, &base, &ncu);
  if (rc) return rc;

  Inst* dinst = reinterpret_cast<Inst*>(base);
  Clock* dclk = reinterpret_cast<Clock*>(base + 1024);
  double* dscratch = reinterpret_cast<double*>(base + kHeadBytes);

  Inst I;
  std::memset(&I, 0, sizeof(I));
  I.nx1 = d.nx1;
  I.nx2 = d.nx2;
  I.is = 4;
  I.ie = d.nx1 + 3;
  I.js = 4;
  I.je = d.nx2 + 3;
  I.sx = sx;
  I.sy = sy;
  I.pitch = pitch;
  I.flux = d.flux;
  for (int b = 0; b < 4; b++) I.bc[b] = bcs[b];
  I.dx1 = (d.x1max - d.x1min) / (double)d.nx1;  // init_grid.c:60
  I.dx2 = (d.x2max - d.x2min) / (double)d.nx2;
  I.gamma = d.gamma;
  I.g1 = d.gamma - 1.0;  // main.c:93-94
  I.g2 = d.gamma - 2.0;
  // state[cur = 0] is read, state[1] receives the integrated state (k_end's
  // swap of the clock's index is not read back)
  I.state[0] = const_cast<double*>(state_in);
  I.state[1] = state_out;
  I.scratch = dscratch;
  I.prims = nullptr;
  I.ps = plane_stride;
  I.clk = dclk;

  BatchLaunch bl{};
  bl.insts = dinst;
  bl.ninst = 1;
  bl.ncu = ncu;
  // the handle's default launch shape for one grid
  bl.shape.x2_seg = default_x2_seg((long)cdiv(d.nx1 + 6, kX2Block) * (d.nx2 + 4));
  bl.shape.persist = 0;
  bl.shape.xcd_remap = 1;
  bl.shape.corr_rows = 2;
  bl.shape.final_rows = 8;
  bl.shape.x2_kind = 1;
  for (int kind = 0; kind < NKINDS; kind++) {
    I.tile0[kind] = 0;
    bl.ntiles[kind] = tiles_of(kind, d.nx1, d.nx2, bl.shape.x2_seg, bl.shape.corr_rows,
                              bl.shape.final_rows);
  }
  bl.flux_mask = 1 << d.flux;

  const hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_static_head, dim3(1), dim3(256), 0, s, I, dclk, dinst, dt, d.cour_no);
  HIP_TRY(hipGetLastError());
  const size_t nclear = (size_t)NSCRATCH * plane_stride;
  const size_t blocks = (nclear + 255) / 256;
  hipLaunchKernelGGL(k_static_clear, dim3(blocks < 8192 ? (unsigned)blocks : 8192u), dim3(256),
                     0, s, dscratch, nclear);
  HIP_TRY(hipGetLastError());
  const LaunchFn launch = (d.variant == ATHMI_FAST) ? athmi_fast_launch : athmi_exact_launch;
  for (int k : athmi_rt::kSeqIntegrate) HIP_TRY(launch(&bl, k, s));
  return ATHMI_OK;
}

}  // extern "C"
