// Strided batched GEMM on CDNA4 matrix cores: BF16 operands, FP32 accumulate.
//
//   C[b](i,j) = alpha * sum_k A[b](i,k) B[b](k,j) + beta * C[b](i,j)  (+ bias, GELU)
//
// Operands live in HBM as FP32 with arbitrary element strides (so transposes
// and the planar complex layouts of the corrector network need no copies).
// A 256-thread workgroup (4 wavefronts) computes a TM x TN tile of C, each
// wavefront one 32 x 32 block with v_mfma_f32_32x32x16_bf16.  The tile shape
// follows the matrix: the corrector's products are skinny (one side is a
// 32-channel or 32-mode dimension), so N <= 32 takes 128 x 32 tiles, M <= 32
// takes 32 x 128 tiles and the rest 64 x 64 -- no wavefront idles on a
// 32-wide side -- and batches of many items that each fit one 32 x 32 block
// take one wavefront per item.  Operand tiles of K = 32 are staged through LDS as bf16 with
// k contiguous: the unit-stride axis is read with 16-byte loads where the
// strides and the base pointer allow it, and the next K tile is loaded into
// registers while the matrix cores work on the current one.  MFMA operands
// come out of LDS as 16-byte reads.  With split_k > 1 the K range is divided
// over blockIdx.z and partial products are added to C with FP32 atomics (C
// must then already hold beta*C).
//
// MFMA 32x32x16 operand layout (lane l, 8 bf16 per operand register set):
//   A: row l%32, k = 8*(l/32) + e      B: column l%32, k = 8*(l/32) + e
//   D: 16 floats, register r holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "namsa_hip.h"

namespace namsa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct GemmArgs {
  const float* A;
  long long sAb, sAi, sAk;
  const float* B;
  long long sBb, sBk, sBj;
  float* C;
  long long sCb, sCi, sCj;
  int M, N, K, batch, split_k;
  float alpha, beta;
  const float* bias;  // bias_mode 1: bias[i]; 2: bias[j]
  int bias_mode;
  int gelu;           // 1: C = gelu(value), pre-activation stored to P
  float* P;           // same strides as C
};

constexpr int kTK = 32;
constexpr int kLdK = kTK + 8;  // LDS row pitch in bf16: 80 bytes keeps every row 16-byte aligned

__device__ __forceinline__ float gelu_tanh(float x) {
  const float c = 0.7978845608028654f;  // sqrt(2/pi)
  return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}

// One operand tile of R rows (i for A, j for B) by kTK k values, element
// (r, k) at base + (r0 + r) * sr + (k0 + k) * sk, valid for r0 + r < rows
// and k0 + k < k_end.  Each thread carries G groups of 4 elements that are
// consecutive along the operand's unit-stride axis (k when kfast).
template <int R>
struct Tile {
  static constexpr int G = R * kTK / 4 / 256;
  float v[G][4];

  __device__ __forceinline__ void load(const float* __restrict__ base, long long sr, long long sk, int r0,
                                       int rows, int k0, int k_end, bool kfast, bool vec) {
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int grp = threadIdx.x + 256 * q;
      int r, k;
      if (kfast) {
        r = grp / (kTK / 4);
        k = (grp % (kTK / 4)) * 4;
      } else {
        k = grp / (R / 4);
        r = (grp % (R / 4)) * 4;
      }
      const int gr = r0 + r, gk = k0 + k;
      if (vec && (kfast ? (gr < rows && gk + 3 < k_end) : (gr + 3 < rows && gk < k_end))) {
        const float4 f = *reinterpret_cast<const float4*>(base + gr * sr + gk * sk);
        v[q][0] = f.x;
        v[q][1] = f.y;
        v[q][2] = f.z;
        v[q][3] = f.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int er = kfast ? gr : gr + e, ek = kfast ? gk + e : gk;
          v[q][e] = (er < rows && ek < k_end) ? base[er * sr + ek * sk] : 0.0f;
        }
      }
    }
  }

  __device__ __forceinline__ void store(__bf16 (*lds)[kLdK], bool kfast) const {
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int grp = threadIdx.x + 256 * q;
      if (kfast) {
        const int r = grp / (kTK / 4), k = (grp % (kTK / 4)) * 4;
        bf16x4 h;
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = __bf16(v[q][e]);
        *reinterpret_cast<bf16x4*>(&lds[r][k]) = h;
      } else {
        const int k = grp / (R / 4), r = (grp % (R / 4)) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) lds[r + e][k] = __bf16(v[q][e]);
      }
    }
  }
};

// 16-byte loads along the unit-stride axis need that stride to be 1, the
// other stride and the batch stride to keep 4-float alignment, and an
// aligned base.
__device__ __forceinline__ bool vec_ok(const float* base, long long s_unit, long long s_other, long long s_batch) {
  return s_unit == 1 && (s_other & 3) == 0 && (s_batch & 3) == 0 &&
         (reinterpret_cast<uintptr_t>(base) & 15) == 0;
}

template <int WM, int WN>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  constexpr int TM = 32 * WM, TN = 32 * WN;
  __shared__ __align__(16) __bf16 As[TM][kLdK];
  __shared__ __align__(16) __bf16 Bs[TN][kLdK];  // transposed: [j][k]

  const int tiles_n = (g.N + TN - 1) / TN;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  const int b = blockIdx.y;
  const int split = blockIdx.z;
  const int i0 = tm * TM, j0 = tn * TN;
  const int k_per = ((g.K + g.split_k - 1) / g.split_k + kTK - 1) / kTK * kTK;
  const int k_begin = split * k_per;
  const int k_end = min(g.K, k_begin + k_per);

  const float* A = g.A + b * g.sAb;
  const float* B = g.B + b * g.sBb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WN, wn = wave % WN;

  // A rows are i (stride sAi), B rows are j (stride sBj); k is the column
  const bool a_kfast = g.sAk == 1 || (g.sAi != 1 && g.sAk <= g.sAi);
  const bool b_kfast = g.sBk == 1 || (g.sBj != 1 && g.sBk <= g.sBj);
  const bool a_vec = a_kfast ? vec_ok(g.A, g.sAk, g.sAi, g.sAb) : vec_ok(g.A, g.sAi, g.sAk, g.sAb);
  const bool b_vec = b_kfast ? vec_ok(g.B, g.sBk, g.sBj, g.sBb) : vec_ok(g.B, g.sBj, g.sBk, g.sBb);

  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;

  Tile<TM> ta;
  Tile<TN> tb;
  if (k_begin < k_end) {
    ta.load(A, g.sAi, g.sAk, i0, g.M, k_begin, k_end, a_kfast, a_vec);
    tb.load(B, g.sBj, g.sBk, j0, g.N, k_begin, k_end, b_kfast, b_vec);
  }
  for (int k0 = k_begin; k0 < k_end; k0 += kTK) {
    ta.store(As, a_kfast);
    tb.store(Bs, b_kfast);
    __syncthreads();
    if (k0 + kTK < k_end) {  // next K tile in flight while the matrix cores work
      ta.load(A, g.sAi, g.sAk, i0, g.M, k0 + kTK, k_end, a_kfast, a_vec);
      tb.load(B, g.sBj, g.sBk, j0, g.N, k0 + kTK, k_end, b_kfast, b_vec);
    }
    const int ra = wm * 32 + (lane & 31), rb = wn * 32 + (lane & 31);
#pragma unroll
    for (int ks = 0; ks < kTK; ks += 16) {
      const int kk = ks + 8 * (lane >> 5);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[ra][kk]);
      const bf16x8 bb = *reinterpret_cast<const bf16x8*>(&Bs[rb][kk]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc, 0, 0, 0);
    }
    __syncthreads();
  }

  float* C = g.C + b * g.sCb;
  float* P = g.P ? g.P + b * g.sCb : nullptr;
  const int j = j0 + wn * 32 + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = i0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= g.M || j >= g.N) continue;
    float* c = C + i * g.sCi + j * g.sCj;
    float v = g.alpha * acc[r];
    if (g.split_k > 1) {
      atomicAdd(c, v);
      continue;
    }
    if (g.beta != 0.0f) v += g.beta * *c;
    if (g.bias_mode == 1) v += g.bias[i];
    if (g.bias_mode == 2) v += g.bias[j];
    if (g.gelu) {
      P[i * g.sCi + j * g.sCj] = v;
      v = gelu_tanh(v);
    }
    *c = v;
  }
}

// Epilogue of one 32 x 32 MFMA block of item b, rows i0 .. i0 + 31:
// C = alpha acc + beta C (+ bias), optionally keeping the pre-activation in P
// and applying GELU.  Register r of lane l holds row 8(r/4) + 4(l/32) + r%4,
// column l%32.
__device__ __forceinline__ void store_block(const GemmArgs& g, long long b, const f32x16& acc, int i0, int lane) {
  float* C = g.C + b * g.sCb;
  float* P = g.P ? g.P + b * g.sCb : nullptr;
  const int j = lane & 31;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = i0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= g.M || j >= g.N) continue;
    float* c = C + i * g.sCi + j * g.sCj;
    float v = g.alpha * acc[r];
    if (g.beta != 0.0f) v += g.beta * *c;
    if (g.bias_mode == 1) v += g.bias[i];
    if (g.bias_mode == 2) v += g.bias[j];
    if (g.gelu) {
      P[i * g.sCi + j * g.sCj] = v;
      v = gelu_tanh(v);
    }
    *c = v;
  }
}

// Batched products whose every matrix fits one MFMA block (M <= 32, N <= 32;
// the corrector's mode-space y transforms, thousands of 32 x 32 x H items):
// one workgroup per batch item, its four wavefronts taking every fourth
// K step of 16 with the operands read straight into the MFMA layout, the
// partial accumulators summed through LDS.  Splitting K shortens each
// wavefront's chain of dependent loads and MFMAs fourfold; one wavefront
// per item left the loads' latency exposed (about 0.8 TB/s on the operand).
__global__ void __launch_bounds__(256) gemm_small_kernel(GemmArgs g) {
  __shared__ float part[3][16][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long b = blockIdx.x;
  const float* A = g.A + b * g.sAb;
  const float* B = g.B + b * g.sBb;
  const int row = lane & 31, kh = 8 * (lane >> 5);
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int k0 = 16 * wave; k0 < g.K; k0 += 64) {
    bf16x8 a, bb;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + kh + e;
      a[e] = __bf16((row < g.M && k < g.K) ? A[row * g.sAi + k * g.sAk] : 0.0f);
      bb[e] = __bf16((row < g.N && k < g.K) ? B[k * g.sBk + row * g.sBj] : 0.0f);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc, 0, 0, 0);
  }
  if (wave > 0) {
#pragma unroll
    for (int r = 0; r < 16; ++r) part[wave - 1][r][lane] = acc[r];
  }
  __syncthreads();
  if (wave > 0) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] += part[0][r][lane] + part[1][r][lane] + part[2][r][lane];
  store_block(g, b, acc, 0, lane);
}

template <int WM, int WN>
int launch_gemm(const GemmArgs& g, hipStream_t st) {
  constexpr int TM = 32 * WM, TN = 32 * WN;
  const long long tiles = static_cast<long long>((g.M + TM - 1) / TM) * ((g.N + TN - 1) / TN);
  if (tiles > 0x7fffffffLL) return fail(kErrArg, "namsa_gemm_bf16: too many tiles");
  hipLaunchKernelGGL((gemm_bf16_kernel<WM, WN>), dim3(static_cast<unsigned>(tiles), g.batch, g.split_k), dim3(256), 0,
                     st, g);
  return check_hip(hipGetLastError(), "gemm_bf16_kernel launch");
}

}  // namespace namsa

using namespace namsa;

extern "C" {

int namsa_gemm_bf16(const float* A, long long sAb, long long sAi, long long sAk, const float* B,
                    long long sBb, long long sBk, long long sBj, float* C, long long sCb,
                    long long sCi, long long sCj, int M, int N, int K, int batch, int split_k,
                    float alpha, float beta, const float* bias, int bias_mode, int gelu, float* P,
                    void* stream) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0 || batch <= 0 || split_k <= 0 || batch > 65535 ||
      split_k > 65535)
    return fail(kErrArg, "namsa_gemm_bf16: bad shape M=%d N=%d K=%d batch=%d split=%d", M, N, K,
                batch, split_k);
  if (split_k > 1 && (bias_mode || gelu))
    return fail(kErrArg, "namsa_gemm_bf16: split-K cannot apply bias or GELU");
  if ((bias_mode && !bias) || (gelu && !P) || bias_mode < 0 || bias_mode > 2)
    return fail(kErrArg, "namsa_gemm_bf16: missing bias or pre-activation buffer");
  GemmArgs g{A, sAb, sAi, sAk, B, sBb, sBk, sBj, C, sCb, sCi, sCj, M, N, K, batch, split_k,
             alpha, beta, bias, bias_mode, gelu, P};
  hipStream_t st = as_stream(stream);
  if (M <= 32 && N <= 32 && split_k == 1 && batch >= 64) {
    hipLaunchKernelGGL(gemm_small_kernel, dim3(static_cast<unsigned>(batch)), dim3(256), 0, st, g);
    return check_hip(hipGetLastError(), "gemm_small_kernel launch");
  }
  if (N <= 32) return launch_gemm<4, 1>(g, st);
  if (M <= 32) return launch_gemm<1, 4>(g, st);
  return launch_gemm<2, 2>(g, st);
}

}  // extern "C"
