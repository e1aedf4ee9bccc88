// transfer.hip -- the host-array boundary of one instance: the caller's Gas
// array (8 doubles per cell, athena.h:43-52) and the three interface-field
// arrays, row-major [sy][sx], against the SoA state planes [sy][pitch].
//
// The caller's bytes cross PCIe unchanged into a device staging area and the
// AoS <-> SoA conversion runs here, one thread per padded plane element, so
// the host does no per-cell work. Pure data movement: the values are copied,
// never recomputed.
#include <hip/hip_runtime.h>

#include "step_kernels.hpp"

namespace athmi {
namespace {

constexpr int kXferBlock = 256;
constexpr unsigned kXferMaxBlocks = 16384;

// stage: gas[n][8] | b1i[n] | b2i[n] | b3i[n] with n = sy * sx. Every
// element of the ps-long planes is written: cells from the stage, the pitch
// padding and the tail of each plane with zero.
__global__ void __launch_bounds__(kXferBlock) k_state_unpack(const double* __restrict__ stage,
                                                             double* __restrict__ st, size_t ps,
                                                             int pitch, int sx, int sy) {
  const size_t n = (size_t)sy * sx;
  for (size_t k = (size_t)blockIdx.x * kXferBlock + threadIdx.x; k < ps;
       k += (size_t)gridDim.x * kXferBlock) {
    const size_t j = k / (size_t)pitch, i = k - j * (size_t)pitch;
    const bool cell = j < (size_t)sy && i < (size_t)sx;
    const size_t h = cell ? j * sx + i : 0;
#pragma unroll
    for (int q = 0; q < 8; q++) st[q * ps + k] = cell ? stage[8 * h + q] : 0.0;
    st[F_B1I * ps + k] = cell ? stage[8 * n + h] : 0.0;
    st[F_B2I * ps + k] = cell ? stage[9 * n + h] : 0.0;
    st[F_B3I * ps + k] = cell ? stage[10 * n + h] : 0.0;
  }
}

// the inverse over the sy * sx cells
__global__ void __launch_bounds__(kXferBlock) k_state_pack(const double* __restrict__ st,
                                                           double* __restrict__ stage, size_t ps,
                                                           int pitch, int sx, int sy) {
  const size_t n = (size_t)sy * sx;
  for (size_t h = (size_t)blockIdx.x * kXferBlock + threadIdx.x; h < n;
       h += (size_t)gridDim.x * kXferBlock) {
    const size_t j = h / (size_t)sx, i = h - j * (size_t)sx;
    const size_t k = j * (size_t)pitch + i;
#pragma unroll
    for (int q = 0; q < 8; q++) stage[8 * h + q] = st[q * ps + k];
    stage[8 * n + h] = st[F_B1I * ps + k];
    stage[9 * n + h] = st[F_B2I * ps + k];
    stage[10 * n + h] = st[F_B3I * ps + k];
  }
}

unsigned blocks_for(size_t n) {
  const size_t b = (n + kXferBlock - 1) / kXferBlock;
  return (unsigned)(b < kXferMaxBlocks ? (b ? b : 1) : kXferMaxBlocks);
}

}  // namespace

hipError_t launch_state_unpack(const double* stage, double* st, size_t ps, int pitch, int sx,
                               int sy, hipStream_t s) {
  hipLaunchKernelGGL(k_state_unpack, dim3(blocks_for(ps)), dim3(kXferBlock), 0, s, stage, st, ps,
                     pitch, sx, sy);
  return hipGetLastError();
}

hipError_t launch_state_pack(const double* st, double* stage, size_t ps, int pitch, int sx,
                             int sy, hipStream_t s) {
  hipLaunchKernelGGL(k_state_pack, dim3(blocks_for((size_t)sy * sx)), dim3(kXferBlock), 0, s,
                     st, stage, ps, pitch, sx, sy);
  return hipGetLastError();
}

}  // namespace athmi
