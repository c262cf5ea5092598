// Strided batched GEMM on CDNA4 matrix cores: BF16 operands, FP32 accumulate.
//
//   C[b](i,j) = alpha * sum_k A[b](i,k) B[b](k,j) + beta * C[b](i,j)  (+ bias, GELU)
//
// Operands live in HBM as FP32 with arbitrary element strides (so transposes
// and the planar complex layouts of the corrector network need no copies);
// a 256-thread workgroup stages a 64x32 tile of A and a 32x64 tile of B
// through LDS as bf16 (k contiguous), and each of its four wavefronts
// accumulates a 32x32 block with v_mfma_f32_32x32x16_bf16.  With split_k > 1
// the K range is divided over blockIdx.z and partial products are added to C
// with FP32 atomics (C must then already hold beta*C).
//
// MFMA 32x32x16 operand layout (lane l, 8 bf16 per operand register set):
//   A: row l%32, k = 8*(l/32) + e      B: column l%32, k = 8*(l/32) + e
//   D: 16 floats, register r holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "namsa_hip.h"

namespace namsa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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

constexpr int kTM = 64, kTN = 64, kTK = 32, kPadK = 8;

__device__ __forceinline__ float gelu_tanh(float x) {
  const float c = 0.7978845608028654f;  // sqrt(2/pi)
  return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}

__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  __shared__ __bf16 As[kTM][kTK + kPadK];
  __shared__ __bf16 Bs[kTN][kTK + kPadK];  // transposed: [j][k]

  const int tiles_n = (g.N + kTN - 1) / kTN;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  const int b = blockIdx.y;
  const int split = blockIdx.z;
  const int i0 = tm * kTM, j0 = tn * kTN;
  const int k_per = ((g.K + g.split_k - 1) / g.split_k + kTK - 1) / kTK * kTK;
  const int k_begin = split * k_per;
  const int k_end = min(g.K, k_begin + k_per);

  const float* A = g.A + b * g.sAb;
  const float* B = g.B + b * g.sBb;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  // staging map: 8 elements per thread per tile, contiguous along the
  // operand's unit-stride axis so global reads coalesce
  const bool a_kfast = g.sAk <= g.sAi;
  const bool b_kfast = g.sBk <= g.sBj;

  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;

  for (int k0 = k_begin; k0 < k_end; k0 += kTK) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + 256 * e;  // 0..2047 over the 64x32 tile
      int ii, kk;
      if (a_kfast) { ii = idx / kTK; kk = idx % kTK; } else { ii = idx % kTM; kk = idx / kTM; }
      const int gi = i0 + ii, gk = k0 + kk;
      float v = 0.0f;
      if (gi < g.M && gk < k_end) v = A[gi * g.sAi + gk * g.sAk];
      As[ii][kk] = __bf16(v);
      int jj, kb;
      if (b_kfast) { jj = idx / kTK; kb = idx % kTK; } else { jj = idx % kTN; kb = idx / kTN; }
      const int gj = j0 + jj, gkb = k0 + kb;
      float w = 0.0f;
      if (gj < g.N && gkb < k_end) w = B[gkb * g.sBk + gj * g.sBj];
      Bs[jj][kb] = __bf16(w);
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kTK; ks += 16) {
      bf16x8 a, bb;
      const int ra = wm * 32 + (lane & 31), rb = wn * 32 + (lane & 31);
      const int kk = ks + 8 * (lane >> 5);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a[e] = As[ra][kk + e];
        bb[e] = Bs[rb][kk + e];
      }
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
  const long long tiles = static_cast<long long>((M + kTM - 1) / kTM) * ((N + kTN - 1) / kTN);
  if (tiles > 0x7fffffffLL) return fail(kErrArg, "namsa_gemm_bf16: too many tiles");
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3(static_cast<unsigned>(tiles), batch, split_k), dim3(256),
                     0, as_stream(stream), g);
  return check_hip(hipGetLastError(), "gemm_bf16_kernel launch");
}

}  // extern "C"
