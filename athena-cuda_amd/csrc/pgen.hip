// pgen.hip -- the problem generators on the device, for many instances in
// one launch: blast.c:25-72 (with the displaced centre of the ensemble),
// field_loop.c:40-137 (iprob = 1) and cpaw2d.c:56-185.
//
// Each thread owns one position (j, i) of [js,je+1] x [is,ie+1] of one
// instance and writes what the host generator writes there; the rest of the
// state buffer is zero, as the reference's calloc'd Grid. Constants that the
// reference takes from libm once per grid (cos/sin of the field angle, the
// wave's direction, sqrt of the box diagonal) are computed on the host with
// the host libm. Built with -ffp-contract=off and correctly rounded sqrt and
// division, blast and field_loop states equal the host generators' bit for
// bit; cpaw2d evaluates sin/cos per cell with the device libm (within an ulp
// of the host's).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "host_problems.hpp"
#include "runtime.hpp"

namespace athmi {
namespace pgenk {

struct PgenInst {
  double* st;  // the instance's current state buffer
  size_t ps;
  int pitch, is, ie, js, je;
  int kind, block0, ntx;  // first block of the instance, blocks along x1
  double dx1, dx2, x1_0, x2_0;
  double c[8];
};

__device__ __forceinline__ void cc_pos(const PgenInst& I, int i, int j, double& x1, double& x2) {
  x1 = I.x1_0 + ((double)(i - I.is) + 0.5) * I.dx1;  // cc_pos.c:25-31, idisp = -is
  x2 = I.x2_0 + ((double)(j - I.js) + 0.5) * I.dx2;
}

// field_loop.c:69-80: the vector potential at the corner (j-1/2, i-1/2)
__device__ __forceinline__ double az_at(const PgenInst& I, int i, int j) {
  double x1c, x2c;
  cc_pos(I, i, j, x1c, x2c);
  const double x1f = x1c - 0.5 * I.dx1, x2f = x2c - 0.5 * I.dx2;
  const double rad = I.c[0], amp = I.c[1];
  if (I.c[6] == 1.0 && (x1f * x1f + x2f * x2f) < rad * rad)
    return amp * (rad - sqrt(x1f * x1f + x2f * x2f));
  return 0.0;
}

// cpaw2d.c:311-315
__device__ __forceinline__ double a3(const PgenInst& I, double x1, double x2) {
  const double sin_a = I.c[0], cos_a = I.c[1], b_par = I.c[2], b_perp = I.c[3], k = I.c[4];
  return b_par * (x1 * sin_a - x2 * cos_a) - (b_perp / k) * cos(k * (x1 * cos_a + x2 * sin_a));
}
// cpaw2d.c:118-131: the face fields at position (j, i)
__device__ __forceinline__ void cpaw_faces(const PgenInst& I, int i, int j, double& b1, double& b2,
                                           double& b3) {
  double x1, x2;
  cc_pos(I, i, j, x1, x2);
  const double cs = cos(I.c[4] * (x1 * I.c[1] + x2 * I.c[0]));
  x1 = x1 - 0.5 * I.dx1;
  x2 = x2 - 0.5 * I.dx2;
  b1 = -(a3(I, x1, x2 + I.dx2) - a3(I, x1, x2)) / I.dx2;
  b2 = (a3(I, x1 + I.dx1, x2) - a3(I, x1, x2)) / I.dx1;
  b3 = I.c[3] * cs;
}

__global__ void __launch_bounds__(kTileX* kTileY) k_pgen(const PgenInst* __restrict__ insts, int n) {
  int lo = 0, hi = n - 1;  // the instance whose block range holds this block
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (insts[mid].block0 <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const PgenInst& I = insts[lo];
  const int local = (int)blockIdx.x - I.block0;
  const int ty = local / I.ntx, tx = local - ty * I.ntx;
  const int i = I.is + tx * kTileX + (int)threadIdx.x;
  const int j = I.js + ty * kTileY + (int)threadIdx.y;
  if (i > I.ie + 1 || j > I.je + 1) return;
  const uint32_t k = (uint32_t)(j * I.pitch + i);
  double* st = I.st;
  const size_t ps = I.ps;
  auto put = [&](int q, uint32_t kk, double v) { st[q * ps + kk] = v; };
  const bool cell = (i <= I.ie && j <= I.je);
  if (I.kind == 0) {  // blast.c:43-70
    if (!cell) return;
    const double bxa = I.c[0], bya = I.c[1];
    put(F_D, k, 1.0);
    put(F_M1, k, 0.0);
    put(F_M2, k, 0.0);
    put(F_M3, k, 0.0);
    put(F_B1I, k, bxa);
    put(F_B2I, k, bya);
    put(F_B3I, k, 0.0);
    put(F_B1C, k, bxa);
    put(F_B2C, k, bya);
    put(F_B3C, k, 0.0);
    if (i == I.ie && I.ie > I.is) put(F_B1I, k + 1, bxa);
    if (j == I.je && I.je > I.js) put(F_B2I, k + I.pitch, bya);
    double x1, x2;
    cc_pos(I, i, j, x1, x2);
    const double rx = x1 - I.c[3], ry = x2 - I.c[4];
    const double rad = sqrt(rx * rx + ry * ry);
    put(F_E, k, (rad < I.c[2]) ? I.c[6] : I.c[5]);
  } else if (I.kind == 1) {  // field_loop.c:84-131
    if (!cell) return;
    const double a00 = az_at(I, i, j), a10 = az_at(I, i, j + 1);
    const double a01 = az_at(I, i + 1, j), a11 = az_at(I, i + 1, j + 1);
    const double b1 = (a10 - a00) / I.dx2, b1p = (a11 - a01) / I.dx2;
    const double b2 = -(a01 - a00) / I.dx1, b2p = -(a11 - a10) / I.dx1;
    put(F_D, k, 1.0);
    put(F_M1, k, I.c[2]);
    put(F_M2, k, I.c[3]);
    put(F_M3, k, 0.0);
    put(F_B1I, k, b1);
    put(F_B2I, k, b2);
    put(F_B3I, k, 0.0);
    if (i == I.ie) put(F_B1I, k + 1, b1p);
    if (j == I.je) put(F_B2I, k + I.pitch, b2p);
    const double b1c = 0.5 * (b1 + b1p), b2c = 0.5 * (b2 + b2p), b3c = 0.0;
    put(F_B1C, k, b1c);
    put(F_B2C, k, b2c);
    put(F_B3C, k, b3c);
    put(F_E, k, I.c[4] + 0.5 * (b1c * b1c + b2c * b2c + b3c * b3c) + I.c[5]);
  } else {  // cpaw2d.c:113-185
    double b1, b2, b3;
    cpaw_faces(I, i, j, b1, b2, b3);
    put(F_B1I, k, b1);
    put(F_B2I, k, b2);
    put(F_B3I, k, b3);
    if (!cell) return;
    double b1p, b2p, d0, d1;
    cpaw_faces(I, i + 1, j, b1p, d0, d1);
    cpaw_faces(I, i, j + 1, d0, b2p, d1);
    double x1, x2;
    cc_pos(I, i, j, x1, x2);
    const double sin_a = I.c[0], cos_a = I.c[1], b_perp = I.c[3], kp = I.c[4];
    const double v_par = I.c[5], v_perp = I.c[6];
    const double sn = sin(kp * (x1 * cos_a + x2 * sin_a));
    const double cs = cos(kp * (x1 * cos_a + x2 * sin_a));
    const double den = 1.0;
    const double m1 = den * (v_par * cos_a - v_perp * sn * sin_a);
    const double m2 = den * (v_par * sin_a + v_perp * sn * cos_a);
    const double m3 = den * v_perp * cs;
    const double b1c = 0.5 * (b1 + b1p), b2c = 0.5 * (b2 + b2p), b3c = b_perp * cs;
    put(F_D, k, den);
    put(F_M1, k, m1);
    put(F_M2, k, m2);
    put(F_M3, k, m3);
    put(F_B1C, k, b1c);
    put(F_B2C, k, b2c);
    put(F_B3C, k, b3c);
    put(F_E, k, I.c[7] + 0.5 * (b1c * b1c + b2c * b2c + b3c * b3c) +
                    0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / den);
  }
}

}  // namespace pgenk
}  // namespace athmi

using namespace athmi::pgenk;

extern "C" int athmi_pgen_device(athmi_grid* g, int first, int n, const athmi_pgen_params* params,
                                 double* lambda) {
  if (!g || !params || n <= 0) return fail(ATHMI_EINVAL, "athmi_pgen_device: bad arguments");
  if (first < 0 || first + n > g->n)
    return fail(ATHMI_EINVAL, "athmi_pgen_device: instance range out of bounds");
  constexpr double kPi = 3.14159265358979323846;  // defs.h:99
  std::vector<PgenInst> hs(n);
  int blocks = 0;
  for (int m = 0; m < n; m++) {
    const int inst = first + m;
    const athmi_pgen_params& P = params[m];
    if (P.problem < ATHMI_PGEN_BLAST || P.problem > ATHMI_PGEN_CPAW2D)
      return fail(ATHMI_EINVAL, "athmi_pgen_device: unknown problem");
    Clock c;
    int rc = g->read_clock(inst, c);
    if (rc) return rc;
    const Inst& I = g->hinst[inst];
    const HostGeom G = geom_of(g, inst);
    PgenInst& S = hs[m];
    std::memset(&S, 0, sizeof S);
    S.st = I.state[c.cur];
    S.ps = I.ps;
    S.pitch = I.pitch;
    S.is = I.is;
    S.ie = I.ie;
    S.js = I.js;
    S.je = I.je;
    S.kind = P.problem;
    S.dx1 = G.dx1;
    S.dx2 = G.dx2;
    S.x1_0 = G.x1_0;
    S.x2_0 = G.x2_0;
    S.ntx = cdiv(I.nx1 + 1, kTileX);
    S.block0 = blocks;
    blocks += S.ntx * cdiv(I.nx2 + 1, kTileY);
    const double* p = P.p;
    if (lambda) lambda[m] = 0.0;
    if (P.problem == ATHMI_PGEN_BLAST) {  // radius, pamb, prat, b0, angle, cx, cy
      const double theta = (kPi / 180.0) * p[4];
      const double bxa = p[3] * std::cos(theta), bya = p[3] * std::sin(theta), bza = 0.0;
      const double da = 1.0, ua = 0.0, va = 0.0, wa = 0.0;
      const double e_mag = 0.5 * (bxa * bxa + bya * bya + bza * bza);
      const double e_kin = 0.5 * da * (ua * ua + va * va + wa * wa);
      const double p_in = p[2] * p[1];
      S.c[0] = bxa;
      S.c[1] = bya;
      S.c[2] = p[0];
      S.c[3] = p[5];
      S.c[4] = p[6];
      S.c[5] = p[1] / G.gamma_1 + e_mag + e_kin;
      S.c[6] = p_in / G.gamma_1 + e_mag + e_kin;
    } else if (P.problem == ATHMI_PGEN_FIELD_LOOP) {  // rad, amp, vflow, iprob
      const double x1size = G.dx1 * (double)G.Nx1, x2size = G.dx2 * (double)G.Nx2;
      const double diag = std::sqrt(x1size * x1size + x2size * x2size);
      const double d = 1.0, m1 = d * p[2] * x1size / diag, m2 = d * p[2] * x2size / diag, m3 = 0.0;
      S.c[0] = p[0];
      S.c[1] = p[1];
      S.c[2] = m1;
      S.c[3] = m2;
      S.c[4] = 1.0 / G.gamma_1;
      S.c[5] = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / d;
      S.c[6] = (double)(int)p[3];
    } else {  // angle, b_par, b_perp, pres, v_par
      const CpawConst k = cpaw2d_constants(G, p[0], p[1], p[2]);
      S.c[0] = k.sin_a;
      S.c[1] = k.cos_a;
      S.c[2] = k.b_par;
      S.c[3] = k.b_perp;
      S.c[4] = k.k_par;
      S.c[5] = p[4];
      S.c[6] = p[2] / std::sqrt(1.0);
      S.c[7] = p[3] / G.gamma_1;
      if (lambda) lambda[m] = k.lambda;
    }
    HIP_TRY(hipMemsetAsync(I.state[c.cur], 0, NSTATE * I.ps * sizeof(double), g->stream));
  }
  PgenInst* dev = nullptr;
  HIP_TRY(hipMalloc(&dev, sizeof(PgenInst) * n));
  hipError_t e = hipMemcpyAsync(dev, hs.data(), sizeof(PgenInst) * n, hipMemcpyHostToDevice,
                                g->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pgen, dim3((unsigned)blocks), dim3(kTileX, kTileY), 0, g->stream, dev, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return fail(ATHMI_EHIP, std::string("athmi_pgen_device: ") + hipGetErrorString(e));
  // cpaw2d keeps its initial state for the error norm (cpaw2d.c:222-302)
  for (int m = 0; m < n; m++) {
    if (params[m].problem != ATHMI_PGEN_CPAW2D) {
      g->cpaw_soln[first + m].clear();
      continue;
    }
    const int inst = first + m;
    const HostGeom G = geom_of(g, inst);
    HostState S(G);
    int rc = athmi_download(g, inst, S.gas.data(), S.b1i.data(), S.b2i.data(), S.b3i.data());
    if (rc) return rc;
    const size_t np = (size_t)G.sx * G.sy;
    std::vector<double>& soln = g->cpaw_soln[inst];
    soln.assign(8 * np, 0.0);
    for (int j = G.js; j <= G.je; j++)
      for (int i = G.is; i <= G.ie; i++) {
        const size_t kk = (size_t)j * G.sx + i;
        for (int q = 0; q < 8; q++) soln[q * np + kk] = S.gas[8 * kk + q];
      }
  }
  return ATHMI_OK;
}
