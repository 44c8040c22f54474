// cc_mfma.hip — the k = 2 rebuild of G on the matrix cores (v_mfma_i32_32x32x32_i8), for the many start
// partitions of a Neighborhood batch.
//
// With z[s][u] = [label_s(u) == 0] and the adjacency A as 0/1 bytes, G[s][0][v] = sum_u z[s][u] A[v][u] is an
// integer matrix product (batch x n) . (n x n)^T, exact in int32; G[s][1][v] = deg(v) - G[s][0][v] (the
// reference's KernelRebuild2 identity, cuda/cc_population_kernels.cuh:94-136). A is expanded once per solve into
// an n_pad x n_pad byte matrix (n_pad = n rounded up to 128, zero padding); z is expanded per batch from the slot
// labels into batch_pad x n_pad bytes. The sums intra[s] = sum_v G[s][label v][v] that the objective needs are
// accumulated in the same kernel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cc_internal.hpp"

namespace cc::hip {
namespace {

using V4i = int __attribute__((ext_vector_type(4)));
using V16i = int __attribute__((ext_vector_type(16)));

constexpr int kTile = 64;  // output tile: 64 starts x 64 vertices per workgroup, one 32 x 32 block per wave
constexpr int kMfmaSlab = 256;  // bytes of K staged per step; rows and columns are padded to whole slabs

// Bytes of one 64-bit adjacency word: byte i = bit i.
__device__ __forceinline__ void ExpandWord(uint64_t word, uint4* out) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t part[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const uint32_t bits = static_cast<uint32_t>(word >> (16 * q + 4 * d)) & 0xFu;
      // spread 4 bits to the low bit of 4 bytes
      part[d] = (bits & 1u) | ((bits & 2u) << 7) | ((bits & 4u) << 14) | ((bits & 8u) << 21);
    }
    out[q] = make_uint4(part[0], part[1], part[2], part[3]);
  }
}

// adj8[v][u] for v, u < n_pad; one thread per 64-bit word of the padded matrix.
__global__ void __launch_bounds__(256) KernelExpandAdjacency(const uint64_t* bits, int n, int words, int n_pad,
                                                            uint4* adj8) {
  const int row = blockIdx.y;
  const int w = blockIdx.x * 256 + threadIdx.x;
  const int row_words = n_pad / 64;
  if (w >= row_words) return;
  const uint64_t word = (row < n && w < words) ? bits[static_cast<size_t>(row) * words + w] : 0ull;
  ExpandWord(word, adj8 + (static_cast<size_t>(row) * row_words + w) * 4);
}

// z[s][u] = [labels[s][u] == 0] for s < count, u < n; zero elsewhere in the batch_pad x n_pad block.
__global__ void __launch_bounds__(256) KernelStartMatrix(const RunDev* run, int count, int n_pad, uint32_t* z) {
  const RunDev& r = *run;
  const int s = blockIdx.y;
  const int q = blockIdx.x * 256 + threadIdx.x;  // 4-byte group
  if (q * 4 >= n_pad) return;
  uint32_t packed = 0;
  if (s < count) {
    const uint8_t* labels = r.pop.labels + static_cast<size_t>(s) * r.n;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int u = q * 4 + b;
      if (u < r.n && labels[u] == 0) packed |= 1u << (8 * b);
    }
  }
  z[(static_cast<size_t>(s) * n_pad) / 4 + q] = packed;
}

// One workgroup: 64 starts x 64 vertices; wave w computes starts s0 + 32 (w / 2) .. + 31 and vertices
// v0 + 32 (w % 2) .. + 31 with one 32 x 32 int32 accumulator. Operand lanes: row (start or vertex) = lane % 32, 16
// consecutive k from 16 (lane / 32) within each 32-wide k step, the same k assignment for both operands; the
// accumulator element j of a lane is start 8 (j / 4) + 4 (lane / 32) + j % 4 of the block, vertex lane % 32.
__global__ void __launch_bounds__(256) KernelRebuildMfmaK2(const RunDev* run, int count, int n_pad, const int8_t* adj8,
                                                          const int8_t* z) {
  const RunDev& r = *run;
  const int n = r.n;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int s0 = blockIdx.y * kTile, v0 = blockIdx.x * kTile;
  const int s_blk = s0 + 32 * (wave >> 1);
  const int v_blk = v0 + 32 * (wave & 1);
  const int row = lane & 31;
  const int kq = 16 * (lane >> 5);

  // K slabs of kSlab bytes of the 64 z rows and 64 adjacency rows of the tile, staged through LDS with coalesced
  // 16-byte loads (16 consecutive threads per 256-byte row segment); the next slab is loaded into registers while
  // the current one feeds the matrix cores. Rows are padded by 16 bytes in LDS against bank conflicts.
  constexpr int kSlab = kMfmaSlab;
  constexpr int kPitch = kSlab + 16;
  constexpr int kChunks = kTile * kSlab / 16 / 256;  // 16-byte chunks per thread per matrix
  __shared__ __align__(16) int8_t lz[kTile * kPitch];
  __shared__ __align__(16) int8_t la[kTile * kPitch];
  const int t = threadIdx.x;
  // thread t moves chunks t + 256 i of each matrix: row c / (kSlab / 16), 16-byte column c % (kSlab / 16)
  constexpr int kPerRow = kSlab / 16;
  const int col = (t % kPerRow) * 16;
  // Two register sets: slab i + 2 is loaded while slab i is staged and multiplied, so two slabs of loads are in
  // flight behind the matrix cores.
  int rows[kChunks];
#pragma unroll
  for (int i = 0; i < kChunks; ++i) rows[i] = (t + 256 * i) / kPerRow;
  auto load = [&](V4i (&pz)[kChunks], V4i (&pa)[kChunks], int k) {
#pragma unroll
    for (int i = 0; i < kChunks; ++i) {
      pz[i] = *reinterpret_cast<const V4i*>(z + static_cast<size_t>(s0 + rows[i]) * n_pad + k + col);
      pa[i] = *reinterpret_cast<const V4i*>(adj8 + static_cast<size_t>(v0 + rows[i]) * n_pad + k + col);
    }
  };
  const int8_t* rz = lz + (32 * (wave >> 1) + row) * kPitch + kq;
  const int8_t* ra = la + (32 * (wave & 1) + row) * kPitch + kq;
  V16i acc = {};
  auto step = [&](V4i (&pz)[kChunks], V4i (&pa)[kChunks], int k0) {
    __syncthreads();  // every wave is done with the previous slab
#pragma unroll
    for (int i = 0; i < kChunks; ++i) {
      *reinterpret_cast<V4i*>(lz + rows[i] * kPitch + col) = pz[i];
      *reinterpret_cast<V4i*>(la + rows[i] * kPitch + col) = pa[i];
    }
    __syncthreads();
    if (k0 + 2 * kSlab < n_pad) load(pz, pa, k0 + 2 * kSlab);  // uniform
#pragma unroll
    for (int kk = 0; kk < kSlab; kk += 32) {
      const V4i a = *reinterpret_cast<const V4i*>(rz + kk);
      const V4i b = *reinterpret_cast<const V4i*>(ra + kk);
      acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc, 0, 0, 0);
    }
  };
  V4i pz0[kChunks], pa0[kChunks], pz1[kChunks], pa1[kChunks];
  load(pz0, pa0, 0);
  if (kSlab < n_pad) load(pz1, pa1, kSlab);
  for (int k0 = 0; k0 < n_pad; k0 += 2 * kSlab) {
    step(pz0, pa0, k0);
    if (k0 + kSlab < n_pad) step(pz1, pa1, k0 + kSlab);
  }

  const int v = v_blk + row;
  const int degree = v < n ? r.deg[v] : 0;
  int16_t* g = static_cast<int16_t*>(r.pop.g);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int s = s_blk + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    int diag = 0;
    if (s < count && v < n) {
      const int g0 = acc[j];
      const int g1 = degree - g0;
      g[(static_cast<size_t>(s) * 2) * n + v] = static_cast<int16_t>(g0);
      g[(static_cast<size_t>(s) * 2 + 1) * n + v] = static_cast<int16_t>(g1);
      diag = z[static_cast<size_t>(s) * n_pad + v] ? g0 : g1;
    }
    // sum over the 32 vertices of this half-wave (lanes 0-31 and 32-63 hold different starts)
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) diag += __shfl_xor(diag, off, 32);
    if ((lane & 31) == 0 && s < count && diag) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&r.intra[s]), static_cast<unsigned long long>(diag));
    }
  }
}

}  // namespace

// Rows and columns padded to whole K slabs (hence also to whole 64-row tiles).
size_t MfmaPad(int n) { return (static_cast<size_t>(n) + kMfmaSlab - 1) / kMfmaSlab * kMfmaSlab; }

void LaunchExpandAdjacency(const uint64_t* bits, int n, int words, int n_pad, int8_t* adj8, hipStream_t stream) {
  const int row_words = n_pad / 64;
  const dim3 grid((row_words + 255) / 256, n_pad);
  KernelExpandAdjacency<<<grid, 256, 0, stream>>>(bits, n, words, n_pad, reinterpret_cast<uint4*>(adj8));
}

void LaunchRebuildMfmaK2(const RunDev* run, int count, int n_pad, const int8_t* adj8, int8_t* z, hipStream_t stream) {
  const int count_pad = static_cast<int>(MfmaPad(count));
  KernelStartMatrix<<<dim3((n_pad / 4 + 255) / 256, count_pad), 256, 0, stream>>>(run, count, n_pad,
                                                                                   reinterpret_cast<uint32_t*>(z));
  KernelRebuildMfmaK2<<<dim3(n_pad / kTile, count_pad / kTile), 256, 0, stream>>>(run, count, n_pad, adj8, z);
}

}  // namespace cc::hip
