// Training-record encoders for 4D-STEM datasets (the step after the slice
// loop in the reference's dataset scripts, scripts/summit_scripts/utils.py:74-118).
//
// namsa_cbed_records: process_cbed (utils.py:104-118). Each CBED [Ny][Nx] is
// standardised by its own mean and population standard deviation and stored
// as IEEE half. Two launches per batch: rec_stats_kernel writes per-chunk FP64
// partial sums, shifted by the item's first value to keep the one-pass
// variance well conditioned. rec_standardize_kernel reduces an item's partials,
// then writes (x - mean) / std with mean and std rounded to float32 and a
// correctly rounded float32 subtraction and division (the float32 arithmetic
// numpy applies to a float32 array), converted to half round-to-nearest-even.
// Both passes cover every item in one launch (up to 65535 items per launch).
// Batches small enough to keep the second read in the MALL or the L2 measured
// slower on MI355X (12-96 MB: 4.12-2.11 ms against 1.91 ms for 4096 items of
// 512^2), because the launch count grows faster than the reads shrink.
//
// namsa_cbed_records_from_moments: the same records from each item's sum and
// sum of squares (unshifted), as the CBED readout of namsa_multislice_moments
// keeps them; only rec_standardize_kernel runs, so the CBEDs are read once.
//
// namsa_potential_record: process_potential (utils.py:74-102): the mean over
// slices of imag(t), a Gaussian filter (sigma 1.2, truncate 4, 'reflect'
// edges, axis 0 then axis 1, float32 between the passes, as
// scipy.ndimage.gaussian_filter computes it), the crop
// [Ny/4, 3Nx/4) on both axes, then standardisation with the deviation floored
// at 1/sqrt(size), to half. Only the cropped rows and columns are filtered.
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "namsa_hip.h"

namespace namsa {

constexpr int kRecThreads = 256;
constexpr long long kRecChunk = 16384;      // elements per workgroup
constexpr int kRecUnroll = 4;               // float4 loads in flight per thread
constexpr int kMaxRadius = 31;

struct GaussWeights {
  double w[kMaxRadius + 1];                  // w[0] centre, w[j] = weight at distance j
  int radius;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of a pair over the workgroup; the result is valid in thread 0.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
  __shared__ double red[2][kRecThreads / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0.0;
    b = 0.0;
    for (int i = 0; i < kRecThreads / 64; ++i) {
      a += red[0][i];
      b += red[1][i];
    }
  }
}

// grid (chunks, items): partial sums of (x - x0) and (x - x0)^2, x0 = the item's first value.
template <bool VEC>
__global__ __launch_bounds__(kRecThreads) void rec_stats_kernel(const float* __restrict__ x, long long n, int chunks,
                                                                double* __restrict__ part) {
  const int c = blockIdx.x;
  const long long item = blockIdx.y;
  const float* xi = x + item * n;
  const double x0 = xi[0];
  const long long b = c * kRecChunk, e = (b + kRecChunk < n) ? b + kRecChunk : n;
  double s = 0.0, ss = 0.0;
  if constexpr (VEC) {
    const float4* v4 = reinterpret_cast<const float4*>(xi);
    const long long e4 = e >> 2;
    long long i = (b >> 2) + threadIdx.x;
    auto acc = [&](const float4& v) {
      const double d0 = double(v.x) - x0, d1 = double(v.y) - x0, d2 = double(v.z) - x0, d3 = double(v.w) - x0;
      s += (d0 + d1) + (d2 + d3);
      ss += (d0 * d0 + d1 * d1) + (d2 * d2 + d3 * d3);
    };
    for (; i + (kRecUnroll - 1) * kRecThreads < e4; i += kRecUnroll * kRecThreads) {
      float4 v[kRecUnroll];                     // all loads issued before the FP64 sums
#pragma unroll
      for (int u = 0; u < kRecUnroll; ++u) v[u] = v4[i + u * kRecThreads];
#pragma unroll
      for (int u = 0; u < kRecUnroll; ++u) acc(v[u]);
    }
    for (; i < e4; i += kRecThreads) acc(v4[i]);
  } else {
    for (long long i = b + threadIdx.x; i < e; i += kRecThreads) {
      const double d = double(xi[i]) - x0;
      s += d;
      ss += d * d;
    }
  }
  block_sum2(s, ss);
  if (threadIdx.x == 0) {
    part[(item * chunks + c) * 2] = s;
    part[(item * chunks + c) * 2 + 1] = ss;
  }
}

__device__ __forceinline__ unsigned short to_half_bits(float y) { return __half_as_ushort(__float2half_rn(y)); }

// grid (chunks, items): reduce the item's nparts partial sums, then standardise the chunk to half.
// shifted: the sums are of (x - x0) with x0 the item's first value (rec_stats_kernel); otherwise of x.
template <bool VEC>
__global__ __launch_bounds__(kRecThreads) void rec_standardize_kernel(const float* __restrict__ x, long long n,
                                                                      int nparts, const double* __restrict__ part,
                                                                      int shifted, double floor_std,
                                                                      unsigned short* __restrict__ out) {
  __shared__ float stat[2];
  const int c = blockIdx.x;
  const long long item = blockIdx.y;
  const float* xi = x + item * n;
  unsigned short* oi = out + item * n;
  double s = 0.0, ss = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kRecThreads) {
    s += part[(item * nparts + i) * 2];
    ss += part[(item * nparts + i) * 2 + 1];
  }
  block_sum2(s, ss);
  if (threadIdx.x == 0) {
    const double dn = double(n);
    const double m = s / dn;
    const double var = ss / dn - m * m;
    const float mean = float((shifted ? double(xi[0]) : 0.0) + m);
    float sd = float(sqrt(var > 0.0 ? var : 0.0));
    if (double(sd) < floor_std) sd = float(floor_std);
    stat[0] = mean;
    stat[1] = sd;
  }
  __syncthreads();
  const float mean = stat[0], sd = stat[1];
  const long long b = c * kRecChunk, e = (b + kRecChunk < n) ? b + kRecChunk : n;
  if constexpr (VEC) {
    const float4* v4 = reinterpret_cast<const float4*>(xi);
    uint2* o4 = reinterpret_cast<uint2*>(oi);
    const long long e4 = e >> 2;
    long long i = (b >> 2) + threadIdx.x;
    auto enc = [&](const float4& v) {
      const unsigned h0 = to_half_bits(__fdiv_rn(v.x - mean, sd)), h1 = to_half_bits(__fdiv_rn(v.y - mean, sd));
      const unsigned h2 = to_half_bits(__fdiv_rn(v.z - mean, sd)), h3 = to_half_bits(__fdiv_rn(v.w - mean, sd));
      return make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
    };
    for (; i + (kRecUnroll - 1) * kRecThreads < e4; i += kRecUnroll * kRecThreads) {
      float4 v[kRecUnroll];
#pragma unroll
      for (int u = 0; u < kRecUnroll; ++u) v[u] = v4[i + u * kRecThreads];
#pragma unroll
      for (int u = 0; u < kRecUnroll; ++u) o4[i + u * kRecThreads] = enc(v[u]);
    }
    for (; i < e4; i += kRecThreads) o4[i] = enc(v4[i]);
  } else {
    for (long long i = b + threadIdx.x; i < e; i += kRecThreads) oi[i] = to_half_bits(__fdiv_rn(xi[i] - mean, sd));
  }
}

// Mean over S slices of imag(t), FP64 sum, float32 result.
__global__ __launch_bounds__(kRecThreads) void rec_slice_mean_kernel(const float2* __restrict__ t, int S,
                                                                     long long plane, float* __restrict__ proj) {
  const long long i = static_cast<long long>(blockIdx.x) * kRecThreads + threadIdx.x;
  if (i >= plane) return;
  double acc = 0.0;
  for (int s = 0; s < S; ++s) acc += double(t[s * plane + i].y);
  proj[i] = float(acc / double(S));
}

// scipy.ndimage 'reflect' (half-sample symmetric) index into [0, n).
__device__ __forceinline__ int reflect_index(int i, int n) {
  const int p = 2 * n;
  int m = i % p;
  if (m < 0) m += p;
  return m < n ? m : p - 1 - m;
}

#pragma clang fp contract(off)
// Axis 0: g1[r][x] = sum_k w_k proj[reflect(r0 + r + k)][x] for the cropped rows, every column.
__global__ __launch_bounds__(kRecThreads) void rec_gauss_rows_kernel(const float* __restrict__ proj, int Ny, int Nx,
                                                                     int r0, int rows, GaussWeights g,
                                                                     float* __restrict__ g1) {
  const long long i = static_cast<long long>(blockIdx.x) * kRecThreads + threadIdx.x;
  if (i >= static_cast<long long>(rows) * Nx) return;
  const int r = static_cast<int>(i / Nx), x = static_cast<int>(i % Nx);
  const int y = r0 + r;
  double acc = double(proj[static_cast<long long>(y) * Nx + x]) * g.w[0];
  for (int j = g.radius; j >= 1; --j) {         // the symmetric-weight order of scipy's correlate1d
    const double lo = proj[static_cast<long long>(reflect_index(y - j, Ny)) * Nx + x];
    const double hi = proj[static_cast<long long>(reflect_index(y + j, Ny)) * Nx + x];
    acc += (lo + hi) * g.w[j];
  }
  g1[i] = float(acc);
}

// Axis 1 on the cropped rows: g2[r][c] = sum_k w_k g1[r][reflect(c0 + c + k)].
__global__ __launch_bounds__(kRecThreads) void rec_gauss_cols_kernel(const float* __restrict__ g1, int rows, int Nx,
                                                                     int c0, int cols, GaussWeights g,
                                                                     float* __restrict__ g2) {
  const long long i = static_cast<long long>(blockIdx.x) * kRecThreads + threadIdx.x;
  if (i >= static_cast<long long>(rows) * cols) return;
  const int r = static_cast<int>(i / cols), cc = static_cast<int>(i % cols);
  const float* line = g1 + static_cast<long long>(r) * Nx;
  const int x = c0 + cc;
  double acc = double(line[x]) * g.w[0];
  for (int j = g.radius; j >= 1; --j) {
    const double lo = line[reflect_index(x - j, Nx)];
    const double hi = line[reflect_index(x + j, Nx)];
    acc += (lo + hi) * g.w[j];
  }
  g2[i] = float(acc);
}
#pragma clang fp contract(on)

// Items per launch: the grid's y limit.
long long record_batch(long long) { return 65535; }

// Standardise `items` items of n floats to half, batch by batch; part holds 2 * chunks * batch doubles.
int standardize_items(const float* x, long long items, long long n, double floor_std, unsigned short* out,
                      double* part, hipStream_t st, const char* who) {
  const long long chunks = (n + kRecChunk - 1) / kRecChunk;
  const bool vec = (n % 4 == 0) && (reinterpret_cast<uintptr_t>(x) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(out) % 8 == 0);
  const long long batch = record_batch(n);
  for (long long b0 = 0; b0 < items; b0 += batch) {
    const long long nb = (items - b0 < batch) ? items - b0 : batch;
    const dim3 grid(static_cast<unsigned>(chunks), static_cast<unsigned>(nb));
    const float* xb = x + b0 * n;
    unsigned short* ob = out + b0 * n;
    if (vec) {
      hipLaunchKernelGGL(rec_stats_kernel<true>, grid, dim3(kRecThreads), 0, st, xb, n, int(chunks), part);
      hipLaunchKernelGGL(rec_standardize_kernel<true>, grid, dim3(kRecThreads), 0, st, xb, n, int(chunks), part, 1,
                         floor_std, ob);
    } else {
      hipLaunchKernelGGL(rec_stats_kernel<false>, grid, dim3(kRecThreads), 0, st, xb, n, int(chunks), part);
      hipLaunchKernelGGL(rec_standardize_kernel<false>, grid, dim3(kRecThreads), 0, st, xb, n, int(chunks), part,
                         1, floor_std, ob);
    }
    const int rc = check_hip(hipGetLastError(), who);
    if (rc != kOk) return rc;
  }
  return kOk;
}

void potential_crop(int Ny, int Nx, int* r0, int* rows, int* cols) {
  const int start = Ny / 4, stop = 3 * Nx / 4;
  *r0 = start;
  *rows = (stop < Ny ? stop : Ny) - start;
  *cols = (stop < Nx ? stop : Nx) - start;
}

}  // namespace namsa

using namespace namsa;

extern "C" {

long long namsa_records_workspace_doubles(int P, int Ny, int Nx) {
  if (P <= 0 || Ny <= 0 || Nx <= 0) return 0;
  const long long n = static_cast<long long>(Ny) * Nx;
  const long long chunks = (n + kRecChunk - 1) / kRecChunk;
  const long long batch = record_batch(n);
  return 2 * chunks * (P < batch ? P : batch);
}

int namsa_cbed_records(const float* cbed, int P, int Ny, int Nx, void* out_f16, double* ws, void* stream) {
  if (!cbed || !out_f16 || !ws || P < 0 || Ny <= 0 || Nx <= 0)
    return fail(kErrArg, "namsa_cbed_records: bad arguments (P=%d Ny=%d Nx=%d)", P, Ny, Nx);
  if (P == 0) return kOk;
  const long long n = static_cast<long long>(Ny) * Nx;
  if ((n + kRecChunk - 1) / kRecChunk > 0x7fffffffLL) return fail(kErrArg, "namsa_cbed_records: plane too large");
  return standardize_items(cbed, P, n, 0.0, static_cast<unsigned short*>(out_f16), ws, as_stream(stream),
                           "namsa_cbed_records launch");
}

int namsa_cbed_records_from_moments(const float* cbed, const double* moments, int P, int Ny, int Nx,
                                    void* out_f16, void* stream) {
  if (!cbed || !moments || !out_f16 || P < 0 || Ny <= 0 || Nx <= 0)
    return fail(kErrArg, "namsa_cbed_records_from_moments: bad arguments (P=%d Ny=%d Nx=%d)", P, Ny, Nx);
  const long long n = static_cast<long long>(Ny) * Nx;
  const long long chunks = (n + kRecChunk - 1) / kRecChunk;
  if (chunks > 0x7fffffffLL) return fail(kErrArg, "namsa_cbed_records_from_moments: plane too large");
  unsigned short* out = static_cast<unsigned short*>(out_f16);
  const bool vec = (n % 4 == 0) && (reinterpret_cast<uintptr_t>(cbed) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(out) % 8 == 0);
  hipStream_t st = as_stream(stream);
  const long long batch = record_batch(n);
  for (long long b0 = 0; b0 < P; b0 += batch) {
    const long long nb = (P - b0 < batch) ? P - b0 : batch;
    const dim3 grid(static_cast<unsigned>(chunks), static_cast<unsigned>(nb));
    if (vec)
      hipLaunchKernelGGL(rec_standardize_kernel<true>, grid, dim3(kRecThreads), 0, st, cbed + b0 * n, n, 1,
                         moments + 2 * b0, 0, 0.0, out + b0 * n);
    else
      hipLaunchKernelGGL(rec_standardize_kernel<false>, grid, dim3(kRecThreads), 0, st, cbed + b0 * n, n, 1,
                         moments + 2 * b0, 0, 0.0, out + b0 * n);
    if (int rc = check_hip(hipGetLastError(), "namsa_cbed_records_from_moments launch")) return rc;
  }
  return kOk;
}

int namsa_potential_record_shape(int Ny, int Nx, int* rows, int* cols) {
  if (Ny <= 0 || Nx <= 0 || !rows || !cols) return fail(kErrArg, "namsa_potential_record_shape: bad arguments");
  int r0;
  potential_crop(Ny, Nx, &r0, rows, cols);
  if (*rows <= 0 || *cols <= 0)
    return fail(kErrArg, "namsa_potential_record_shape: the crop [%d, %d) is empty for %dx%d", Ny / 4, 3 * Nx / 4,
                Ny, Nx);
  return kOk;
}

long long namsa_potential_record_workspace_bytes(int Ny, int Nx) {
  int rows, cols, r0;
  if (Ny <= 0 || Nx <= 0) return 0;
  potential_crop(Ny, Nx, &r0, &rows, &cols);
  if (rows <= 0 || cols <= 0) return 0;
  const long long plane = static_cast<long long>(Ny) * Nx;
  const long long n = static_cast<long long>(rows) * cols;
  const long long chunks = (n + kRecChunk - 1) / kRecChunk;
  auto up = [](long long b) { return (b + 255) / 256 * 256; };
  return up(4 * plane) + up(4ll * rows * Nx) + up(4 * n) + up(16 * chunks);
}

int namsa_potential_record(const void* t, int S, int Ny, int Nx, double sigma, double truncate, void* out_f16,
                           void* ws, void* stream) {
  if (!t || !out_f16 || !ws || S <= 0 || Ny <= 0 || Nx <= 0 || !(sigma > 0) || !(truncate >= 0))
    return fail(kErrArg, "namsa_potential_record: bad arguments (S=%d Ny=%d Nx=%d sigma=%g)", S, Ny, Nx, sigma);
  int r0, rows, cols;
  potential_crop(Ny, Nx, &r0, &rows, &cols);
  if (rows <= 0 || cols <= 0)
    return fail(kErrArg, "namsa_potential_record: the crop [%d, %d) is empty for %dx%d", Ny / 4, 3 * Nx / 4, Ny, Nx);
  // scipy.ndimage._gaussian_kernel1d: radius int(truncate * sigma + 0.5), weights exp(-x^2 / (2 sigma^2)) / sum
  const int radius = static_cast<int>(truncate * sigma + 0.5);
  if (radius > kMaxRadius)
    return fail(kErrArg, "namsa_potential_record: filter radius %d exceeds %d", radius, kMaxRadius);
  GaussWeights g{};
  g.radius = radius;
  double sum = 0.0;
  double full[2 * kMaxRadius + 1];
  for (int k = -radius; k <= radius; ++k) {
    full[k + radius] = std::exp(-0.5 / (sigma * sigma) * double(k) * double(k));
  }
  for (int k = 0; k < 2 * radius + 1; ++k) sum += full[k];
  for (int j = 0; j <= radius; ++j) g.w[j] = full[radius + j] / sum;

  const long long plane = static_cast<long long>(Ny) * Nx;
  const long long n = static_cast<long long>(rows) * cols;
  auto up = [](long long b) { return (b + 255) / 256 * 256; };
  char* base = static_cast<char*>(ws);
  float* proj = reinterpret_cast<float*>(base);
  float* g1 = reinterpret_cast<float*>(base + up(4 * plane));
  float* g2 = reinterpret_cast<float*>(base + up(4 * plane) + up(4ll * rows * Nx));
  double* part = reinterpret_cast<double*>(base + up(4 * plane) + up(4ll * rows * Nx) + up(4 * n));
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(rec_slice_mean_kernel, dim3(static_cast<unsigned>((plane + kRecThreads - 1) / kRecThreads)),
                     dim3(kRecThreads), 0, st, static_cast<const float2*>(t), S, plane, proj);
  hipLaunchKernelGGL(rec_gauss_rows_kernel,
                     dim3(static_cast<unsigned>((static_cast<long long>(rows) * Nx + kRecThreads - 1) / kRecThreads)),
                     dim3(kRecThreads), 0, st, proj, Ny, Nx, r0, rows, g, g1);
  hipLaunchKernelGGL(rec_gauss_cols_kernel, dim3(static_cast<unsigned>((n + kRecThreads - 1) / kRecThreads)),
                     dim3(kRecThreads), 0, st, g1, rows, Nx, r0, cols, g, g2);
  int rc = check_hip(hipGetLastError(), "namsa_potential_record filter launch");
  if (rc != kOk) return rc;
  return standardize_items(g2, 1, n, 1.0 / std::sqrt(double(n)), static_cast<unsigned short*>(out_f16), part, st,
                           "namsa_potential_record standardise launch");
}

}  // extern "C"
