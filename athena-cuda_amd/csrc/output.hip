// output.hip -- the reference's output layer (output.c, dump_binary.c,
// output_ppm.c, dump_tab.c) on a device-resident instance.
//
// The device does the per-cell work: the float conversions of dump_binary,
// the expression, min/max scan and palette lookup of output_ppm. The host
// writes the headers and the bytes. Built with -ffp-contract=off: the float
// and double arithmetic of every converted value is evaluated one rounding
// per operation, as in the reference's host build, so files are identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "athmi_output.h"
#include "palettes.hpp"
#include "runtime.hpp"

namespace athmi {
namespace outk {

// output.c:603-741 expressions and the field_loop user expressions
enum Expr {
  X_D = 0, X_M1, X_M2, X_M3, X_E, X_B1C, X_B2C, X_B3C,  // Gas members = state planes 0..7
  X_ME, X_V1, X_V2, X_V3, X_P, X_CS2, X_S, X_J3, X_BP2
};
static const char* const kExprNames[] = {"d",  "M1", "M2", "M3", "E",   "B1c", "B2c", "B3c", "ME",
                                         "V1", "V2", "V3", "P",  "cs2", "S",   "J3",  "Bp2"};
constexpr int kNumExpr = sizeof(kExprNames) / sizeof(kExprNames[0]);

struct View {
  const double* st;  // current state buffer
  size_t ps;
  int pitch, is, js, nx1, nx2;
  double gamma, g1, dx1, dx2;
};

__device__ __forceinline__ double expr_at(int e, const View& v, uint32_t k) {
  const double* s = v.st;
  const size_t ps = v.ps;
  auto F = [&](int q) { return s[q * ps + k]; };
  switch (e) {
    case X_ME:  // output.c:646-650
      return 0.5 * (F(F_B1C) * F(F_B1C) + F(F_B2C) * F(F_B2C) + F(F_B3C) * F(F_B3C));
    case X_V1:  // output.c:656-664
      return F(F_M1) / F(F_D);
    case X_V2:
      return F(F_M2) / F(F_D);
    case X_V3:
      return F(F_M3) / F(F_D);
    case X_P:  // output.c:666-671
    case X_CS2:
    case X_S: {
      const double x = F(F_E) - 0.5 * (F(F_B1C) * F(F_B1C) + F(F_B2C) * F(F_B2C) +
                                      F(F_B3C) * F(F_B3C)) -
                       0.5 * (F(F_M1) * F(F_M1) + F(F_M2) * F(F_M2) + F(F_M3) * F(F_M3)) / F(F_D);
      if (e == X_P) return v.g1 * x;
      if (e == X_CS2) return v.gamma * v.g1 * x / F(F_D);  // output.c:677-685
      return (v.g1 * x) / pow(F(F_D), v.gamma);            // output.c:693-703
    }
    case X_J3:  // field_loop.c:160-164
      return (F(F_B2I) - s[F_B2I * ps + k - 1]) / v.dx1 -
             (F(F_B1I) - s[F_B1I * ps + k - (uint32_t)v.pitch]) / v.dx2;
    case X_BP2:  // field_loop.c:166-170
      return F(F_B1C) * F(F_B1C) + F(F_B2C) * F(F_B2C);
    default:
      return F(e);
  }
}

// dump_binary.c:121-130: NVAR float planes [n][j][i] of the interior
__global__ void __launch_bounds__(kTileX* kTileY) k_snap(View v, float* out) {
  const int i = (int)blockIdx.x * kTileX + (int)threadIdx.x;
  const int j = (int)blockIdx.y * kTileY + (int)threadIdx.y;
  if (i >= v.nx1 || j >= v.nx2) return;
  const uint32_t k = (uint32_t)((j + v.js) * v.pitch + i + v.is);
  const size_t n = (size_t)v.nx1 * v.nx2, p = (size_t)j * v.nx1 + i;
  for (int q = 0; q < 8; q++) out[q * n + p] = (float)v.st[q * v.ps + k];
}

// subset2 (output.c:558-600): data[j][i] = (float) expr(i + is, j + js); the
// last position holding a NaN goes to *nanpos (it restarts the scan of
// minmax2, utils.c:148-162)
__global__ void __launch_bounds__(kTileX* kTileY) k_expr(View v, int e, float* data, int* nanpos) {
  const int i = (int)blockIdx.x * kTileX + (int)threadIdx.x;
  const int j = (int)blockIdx.y * kTileY + (int)threadIdx.y;
  if (i >= v.nx1 || j >= v.nx2) return;
  const uint32_t k = (uint32_t)((j + v.js) * v.pitch + i + v.is);
  const float x = (float)expr_at(e, v, k);
  const int p = j * v.nx1 + i;
  data[p] = x;
  if (x != x) atomicMax(nanpos, p);
}

__global__ void k_init_nan(int* nanpos) { *nanpos = -1; }

// minmax2's sequential scan (utils.c:148-162, MIN/MAX of defs.h:90-93) keeps
// a later element on ties: over the positions after the last NaN the min is
// the smallest value, the latest one among equals (which tells -0 from +0);
// likewise the max. As a total order on (value, position) it can be reduced
// in any order.
struct Cand {
  float v;
  int p;  // -1: none
};
__device__ __forceinline__ Cand better_min(Cand a, Cand b) {
  if (a.p < 0) return b;
  if (b.p < 0) return a;
  return (b.v < a.v || (b.v == a.v && b.p > a.p)) ? b : a;
}
__device__ __forceinline__ Cand better_max(Cand a, Cand b) {
  if (a.p < 0) return b;
  if (b.p < 0) return a;
  return (b.v > a.v || (b.v == a.v && b.p > a.p)) ? b : a;
}

constexpr int kMmBlock = 256;
constexpr int kMmParts = 512;

__device__ void block_minmax(Cand mn, Cand mx, Cand* out_min, Cand* out_max) {
  __shared__ float smv[kMmBlock], sxv[kMmBlock];
  __shared__ int smp[kMmBlock], sxp[kMmBlock];
  const int t = threadIdx.x;
  smv[t] = mn.v; smp[t] = mn.p; sxv[t] = mx.v; sxp[t] = mx.p;
  __syncthreads();
  for (int off = kMmBlock / 2; off > 0; off >>= 1) {
    if (t < off) {
      const Cand a = better_min(Cand{smv[t], smp[t]}, Cand{smv[t + off], smp[t + off]});
      const Cand b = better_max(Cand{sxv[t], sxp[t]}, Cand{sxv[t + off], sxp[t + off]});
      smv[t] = a.v; smp[t] = a.p; sxv[t] = b.v; sxp[t] = b.p;
    }
    __syncthreads();
  }
  if (t == 0) {
    *out_min = Cand{smv[0], smp[0]};
    *out_max = Cand{sxv[0], sxp[0]};
  }
}

__global__ void __launch_bounds__(kMmBlock) k_minmax_part(const float* data, int n, const int* nanpos,
                                                          Cand* part) {
  Cand mn{0.0f, -1}, mx{0.0f, -1};
  const int first = *nanpos + 1;
  for (int p = first + (int)(blockIdx.x * kMmBlock + threadIdx.x); p < n;
       p += (int)(gridDim.x * kMmBlock)) {
    const Cand c{data[p], p};
    mn = better_min(mn, c);
    mx = better_max(mx, c);
  }
  block_minmax(mn, mx, &part[2 * blockIdx.x], &part[2 * blockIdx.x + 1]);
}

__global__ void __launch_bounds__(kMmBlock) k_minmax_final(const Cand* part, int nparts, Cand* res) {
  Cand mn{0.0f, -1}, mx{0.0f, -1};
  for (int b = threadIdx.x; b < nparts; b += kMmBlock) {
    mn = better_min(mn, part[2 * b]);
    mx = better_max(mx, part[2 * b + 1]);
  }
  block_minmax(mn, mx, &res[0], &res[1]);
}

// compute_rgb (output_ppm.c:144-180) for image row r (= data row nx2-1-r,
// output_ppm.c:117-133)
__global__ void __launch_bounds__(kTileX* kTileY) k_rgb(const float* data, int nx1, int nx2,
                                                        double dmin, double dmax, const float* rgb,
                                                        const float* der, unsigned char* out) {
  const int i = (int)blockIdx.x * kTileX + (int)threadIdx.x;
  const int r = (int)blockIdx.y * kTileY + (int)threadIdx.y;
  if (i >= nx1 || r >= nx2) return;
  const int j = nx2 - 1 - r;
  const float f = data[(size_t)j * nx1 + i];
  const double d = (f != f) ? 0.0 : (double)f;
  int c[3];
  if (dmin == dmax) {
    c[0] = c[1] = c[2] = (d > dmax ? 255 : 0);
  } else {
    const float x = (float)((d - dmin) * 255.0 / (dmax - dmin));
    if (x <= 0.0f || x >= 255.0f) {
      const int e = (x <= 0.0f ? 0 : 255);
      for (int q = 0; q < 3; q++) c[q] = (int)((double)rgb[e * 3 + q] * 255.0);
    } else if (x == x) {
      const int e = (int)x;
      const float w = x - (float)e;
      for (int q = 0; q < 3; q++) c[q] = (int)((double)(rgb[e * 3 + q] + w * der[e * 3 + q]) * 255.0);
    } else {  // a NaN scale (the reference indexes its palette out of range here)
      c[0] = c[1] = c[2] = 0;
    }
  }
  unsigned char* o = out + ((size_t)r * nx1 + i) * 3;
  o[0] = (unsigned char)c[0];
  o[1] = (unsigned char)c[1];
  o[2] = (unsigned char)c[2];
}

}  // namespace outk
}  // namespace athmi

using namespace athmi::outk;

namespace {

int check(athmi_grid* g, int inst, const char* path) {
  if (!g || !path) return fail(ATHMI_EINVAL, "output: null argument");
  if (inst < 0 || inst >= g->n) return fail(ATHMI_EINVAL, "output: instance index out of range");
  return ATHMI_OK;
}

int reserve(athmi_grid* g, size_t bytes) {
  if (bytes <= g->outbuf_bytes) return ATHMI_OK;
  if (g->outbuf) HIP_TRY(hipFree(g->outbuf));
  g->outbuf = nullptr;
  g->outbuf_bytes = 0;
  HIP_TRY(hipMalloc(&g->outbuf, bytes));
  g->outbuf_bytes = bytes;
  return ATHMI_OK;
}

View view_of(athmi_grid* g, int inst, const Clock& c) {
  const Inst& I = g->hinst[inst];
  return View{I.state[c.cur], I.ps, I.pitch, I.is, I.js, I.nx1, I.nx2,
              I.gamma, I.g1, I.dx1, I.dx2};
}

dim3 tiles(int nx1, int nx2) {
  return dim3((unsigned)cdiv(nx1, kTileX), (unsigned)cdiv(nx2, kTileY));
}

// output.c:305-314: the palette's float table (rebuilt from its codes, see
// palettes.hpp) and the forward differences used for interpolation
bool palette_tables(const char* name, std::vector<float>& rgb, std::vector<float>& der) {
  const athmi::pal::Palette* P = nullptr;
  for (int k = 0; k < athmi::pal::kNumPalettes; k++)
    if (std::strcmp(name, athmi::pal::kPalettes[k].name) == 0) P = &athmi::pal::kPalettes[k];
  if (!P) return false;
  rgb.assign(768, 0.0f);
  der.assign(768, 0.0f);
  char buf[64];
  for (int e = 0; e < 768; e++) {
    const int k = P->code[e];
    if (P->jh)
      std::snprintf(buf, sizeof buf, "%.6e", k == 511 ? 255.0 / 256.0 : k / 510.0);
    else
      std::snprintf(buf, sizeof buf, "%.5f", k / 255.0);
    rgb[e] = (float)std::strtod(buf, nullptr);
  }
  for (int q = 0; q < 3; q++)
    for (int i = 0; i < 255; i++) der[3 * i + q] = rgb[3 * (i + 1) + q] - rgb[3 * i + q];
  return true;
}

// a dat_fmt string: one floating conversion of a double ("%12.5e", "%g", ...)
bool valid_dat_fmt(const char* f) {
  if (*f++ != '%') return false;
  while (*f && std::strchr("-+ #0", *f)) f++;
  while (*f >= '0' && *f <= '9') f++;
  if (*f == '.') {
    f++;
    while (*f >= '0' && *f <= '9') f++;
  }
  if (*f == 'l') f++;
  if (!*f || !std::strchr("eEfFgG", *f)) return false;
  return f[1] == 0;
}

struct File {
  FILE* f;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

extern "C" {

int athmi_dump_binary(athmi_grid* g, int inst, const char* path) {
  int rc = check(g, inst, path);
  if (rc) return rc;
  Clock c;
  if ((rc = g->read_clock(inst, c))) return rc;
  const View v = view_of(g, inst, c);
  const size_t n = (size_t)v.nx1 * v.nx2;
  if ((rc = reserve(g, 8 * n * sizeof(float)))) return rc;
  float* dev = (float*)g->outbuf;
  hipLaunchKernelGGL(k_snap, tiles(v.nx1, v.nx2), dim3(kTileX, kTileY), 0, g->stream, v, dev);
  HIP_TRY(hipGetLastError());
  std::vector<float> data(8 * n);
  HIP_TRY(hipMemcpyAsync(data.data(), dev, data.size() * sizeof(float), hipMemcpyDeviceToHost,
                         g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  const HostGeom G = geom_of(g, inst);
  File fp(path, "wb");
  if (!fp.f) return fail(ATHMI_EINVAL, std::string("athmi_dump_binary: cannot open ") + path);
  const int ndata[6] = {G.ie - G.is + 1, G.je - G.js + 1, 0, 8, 0, 0};
  std::fwrite(ndata, sizeof(int), 6, fp.f);
  float dat[2] = {(float)G.gamma_1, 0.0f};
  std::fwrite(dat, sizeof(float), 2, fp.f);
  dat[0] = (float)c.time;
  dat[1] = (float)c.dt;
  std::fwrite(dat, sizeof(float), 2, fp.f);
  std::vector<float> xs(ndata[0]), ys(ndata[1]);
  double x1, x2;
  for (int i = G.is; i <= G.ie; i++) {
    G.cc_pos(i, G.js, &x1, &x2);
    xs[i - G.is] = (float)x1;
  }
  for (int j = G.js; j <= G.je; j++) {
    G.cc_pos(G.is, j, &x1, &x2);
    ys[j - G.js] = (float)x2;
  }
  std::fwrite(xs.data(), sizeof(float), xs.size(), fp.f);
  std::fwrite(ys.data(), sizeof(float), ys.size(), fp.f);
  if (std::fwrite(data.data(), sizeof(float), data.size(), fp.f) != data.size())
    return fail(ATHMI_EINVAL, std::string("athmi_dump_binary: short write to ") + path);
  return ATHMI_OK;
}

int athmi_dump_tab(athmi_grid* g, int inst, const char* path, const char* dat_fmt) {
  int rc = check(g, inst, path);
  if (rc) return rc;
  if (dat_fmt && !valid_dat_fmt(dat_fmt))
    return fail(ATHMI_EINVAL, "athmi_dump_tab: dat_fmt must be one floating conversion");
  Clock c;
  if ((rc = g->read_clock(inst, c))) return rc;
  const Inst& I = g->hinst[inst];
  const HostGeom G = geom_of(g, inst);
  const int il = G.is, iu = G.ie, jl = G.js, ju = G.je;
  const int j = (ju + jl) / 2;  // dump_tab.c:176: the middle row
  std::vector<double> row((size_t)NSTATE * I.sx);
  for (int q = 0; q < NSTATE; q++)
    HIP_TRY(hipMemcpyAsync(row.data() + (size_t)q * I.sx,
                           I.state[c.cur] + q * I.ps + (size_t)j * I.pitch, I.sx * sizeof(double),
                           hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  File fp(path, "w");
  if (!fp.f) return fail(ATHMI_EINVAL, std::string("athmi_dump_tab: cannot open ") + path);
  const std::string fmt = dat_fmt ? std::string(" ") + dat_fmt : std::string(" %12.8e");
  FILE* f = fp.f;
  int col = 1;
  if (G.Nx1 > 1) {
    std::fprintf(f, "# Nx1 = %d\n", iu - il + 1);
    std::fprintf(f, "# x1-size = %g\n", (iu - il + 1) * G.dx1);
  }
  if (G.Nx2 > 1) {
    std::fprintf(f, "# Nx2 = %d\n", ju - jl + 1);
    std::fprintf(f, "# x2-size = %g\n", (ju - jl + 1) * G.dx2);
  }
  std::fprintf(f, "# Time = %g\n", c.time);
  std::fprintf(f, "#\n#");
  if (G.Nx1 > 1) std::fprintf(f, " [%d]=i-zone", col++);
  if (G.Nx2 > 2) std::fprintf(f, " [%d]=j-zone", col++);
  if (G.Nx1 > 1) std::fprintf(f, " [%d]=x1", col++);
  if (G.Nx2 > 2) std::fprintf(f, " [%d]=x2", col++);
  // dump_tab.c:126-160 (the three interface columns are all headed B1i there)
  for (const char* h : {"d", "M1", "M2", "M3", "P", "E", "B1c", "B2c", "B3c", "B1i", "B1i", "B1i"})
    std::fprintf(f, " [%d]=%s", col++, h);
  std::fprintf(f, "\n#\n");
  auto at = [&](int q, int i) { return row[(size_t)q * I.sx + i]; };
  for (int i = il; i <= iu; i++) {
    double x1, x2;
    G.cc_pos(i, j, &x1, &x2);
    if (G.Nx1 > 1) std::fprintf(f, "%d ", i);
    if (G.Nx2 > 1) std::fprintf(f, "%d ", j);
    if (G.Nx1 > 1) std::fprintf(f, fmt.c_str(), x1);
    if (G.Nx2 > 1) std::fprintf(f, fmt.c_str(), x2);
    const double d = at(F_D, i), m1 = at(F_M1, i), m2 = at(F_M2, i), m3 = at(F_M3, i);
    const double b1 = at(F_B1C, i), b2 = at(F_B2C, i), b3 = at(F_B3C, i), e = at(F_E, i);
    for (double val : {d, m1, m2, m3}) std::fprintf(f, fmt.c_str(), val);
    const double KE = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / d;  // dump_tab.c:204-208
    const double ME = 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
    std::fprintf(f, fmt.c_str(), (e - KE - ME) * G.gamma_1);
    std::fprintf(f, fmt.c_str(), e);
    for (double val : {b1, b2, b3, at(F_B1I, i), at(F_B2I, i), at(F_B3I, i)})
      std::fprintf(f, fmt.c_str(), val);
    std::fprintf(f, "\n");
  }
  return ATHMI_OK;
}

int athmi_output_ppm(athmi_grid* g, int inst, const char* path, const athmi_ppm_spec* spec,
                     int num, double* gmin, double* gmax) {
  int rc = check(g, inst, path);
  if (rc) return rc;
  if (!spec || !spec->expr || !spec->palette || !gmin || !gmax)
    return fail(ATHMI_EINVAL, "athmi_output_ppm: null argument");
  int e = -1;
  for (int k = 0; k < kNumExpr; k++)
    if (std::strcmp(spec->expr, kExprNames[k]) == 0) e = k;
  if (e < 0) return fail(ATHMI_EINVAL, std::string("athmi_output_ppm: unknown expression ") + spec->expr);
  std::vector<float> rgb, der;
  if (!palette_tables(spec->palette, rgb, der)) {
    std::string names;
    for (int k = 0; k < athmi::pal::kNumPalettes; k++)
      names += std::string(" ") + athmi::pal::kPalettes[k].name;
    return fail(ATHMI_EINVAL, std::string("athmi_output_ppm: no palette ") + spec->palette +
                                  "; valid names are" + names);
  }
  Clock c;
  if ((rc = g->read_clock(inst, c))) return rc;
  const View v = view_of(g, inst, c);
  const int nx1 = v.nx1, nx2 = v.nx2;
  const size_t n = (size_t)nx1 * nx2;
  // staging: data floats | palette rgb, der | nan position | partials | result | image bytes
  const size_t off_pal = n * sizeof(float);
  const size_t off_nan = off_pal + 2 * 768 * sizeof(float);
  const size_t off_part = round_up(off_nan + sizeof(int), 64);
  const size_t off_res = off_part + 2 * kMmParts * sizeof(Cand);
  const size_t off_img = round_up(off_res + 2 * sizeof(Cand), 64);
  if ((rc = reserve(g, off_img + 3 * n))) return rc;
  char* base = (char*)g->outbuf;
  float* data = (float*)base;
  float* drgb = (float*)(base + off_pal);
  float* dder = drgb + 768;
  int* nanpos = (int*)(base + off_nan);
  Cand* part = (Cand*)(base + off_part);
  Cand* res = (Cand*)(base + off_res);
  unsigned char* img = (unsigned char*)(base + off_img);
  HIP_TRY(hipMemcpyAsync(drgb, rgb.data(), 768 * sizeof(float), hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(dder, der.data(), 768 * sizeof(float), hipMemcpyHostToDevice, g->stream));
  const int nparts = (int)std::min<size_t>(kMmParts, (n + kMmBlock - 1) / kMmBlock);
  hipLaunchKernelGGL(k_init_nan, dim3(1), dim3(1), 0, g->stream, nanpos);
  hipLaunchKernelGGL(k_expr, tiles(nx1, nx2), dim3(kTileX, kTileY), 0, g->stream, v, e, data, nanpos);
  hipLaunchKernelGGL(k_minmax_part, dim3(nparts), dim3(kMmBlock), 0, g->stream, data, (int)n,
                     nanpos, part);
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(kMmBlock), 0, g->stream, part, nparts, res);
  HIP_TRY(hipGetLastError());
  Cand mm[2];
  HIP_TRY(hipMemcpyAsync(mm, res, sizeof mm, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  // minmax2: no position after the last NaN leaves the scan at NaN
  float dmin = mm[0].p < 0 ? NAN : mm[0].v;
  float dmax = mm[1].p < 0 ? NAN : mm[1].v;
  if (num == 0) {  // output_ppm.c:93-100
    *gmin = dmin;
    *gmax = dmax;
  } else {
    *gmin = (dmin < *gmin) ? (double)dmin : *gmin;
    *gmax = (dmax > *gmax) ? (double)dmax : *gmax;
  }
  File fp(path, "w");
  if (!fp.f) return fail(ATHMI_EINVAL, std::string("athmi_output_ppm: cannot open ") + path);
  std::fprintf(fp.f, "P6\n");
  std::fprintf(fp.f, "# dmin = %.7e, dmax = %.7e, gmin = %.7e, gmax = %.7e\n", (double)dmin,
               (double)dmax, *gmin, *gmax);
  std::fprintf(fp.f, "%d %d\n255\n", nx1, nx2);
  if (spec->sdmin != 0) dmin = (float)spec->dmin;  // output_ppm.c:113-114
  if (spec->sdmax != 0) dmax = (float)spec->dmax;
  hipLaunchKernelGGL(k_rgb, tiles(nx1, nx2), dim3(kTileX, kTileY), 0, g->stream, data, nx1, nx2,
                     (double)dmin, (double)dmax, drgb, dder, img);
  HIP_TRY(hipGetLastError());
  std::vector<unsigned char> bytes(3 * n);
  HIP_TRY(hipMemcpyAsync(bytes.data(), img, bytes.size(), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (std::fwrite(bytes.data(), 1, bytes.size(), fp.f) != bytes.size())
    return fail(ATHMI_EINVAL, std::string("athmi_output_ppm: short write to ") + path);
  return ATHMI_OK;
}

}  // extern "C"
