// Projected-potential slice builder for gfx950.
//
// namsa_atomic_tables: Kirkland projected potential per element on the
// reference's oversampled table grid (namsa/msa.py:137-146, :180-196), 2x2 block
// mean (namsa/utils.py:101-108), shifted to a zero minimum, stored as float32.
// K0 is evaluated from its integral form K0(x) = int_0^inf exp(-x cosh t) dt
// (Abramowitz & Stegun 9.6.24) with the trapezoid rule, whose error for this
// doubly-exponentially decaying analytic integrand falls geometrically with
// the step (step <= 0.7/sqrt(x) keeps the relative error near 1e-14).
//
// namsa_build_slices: the reference kernel (namsa/potential_kernels.cu:5-60)
// lets every pixel thread scan every site of its slice, O(N^2 S A).  Here a
// workgroup owns a 16x64 pixel tile of one slice, streams that slice's sites
// through LDS 256 at a time, keeps (in site order, by ballot compaction) only
// the sites whose clipped footprint overlaps the tile, and accumulates table
// values for its pixels in float32 in site order; exp(i sigma V) with the NaN
// guard of potential_kernels.cu:48-57 is fused into the store.
#include "common.hpp"
#include "namsa_hip.h"

namespace namsa {

__device__ double k0_integral(double x) {
  const double h = fmin(0.25, 0.7 / sqrt(x));
  double s = 0.5;
  for (int k = 1; k < 4096; ++k) {
    const double sh = sinh(0.5 * k * h);
    const double a = 2.0 * x * sh * sh;  // x (cosh(kh) - 1)
    if (a > 40.0) break;
    s += exp(-a);
  }
  return exp(-x) * h * s;
}

// Kirkland projected potential (msa.py:181-193), eps = 0.1 A, a0 = 0.529 A, e = 14.4 V A.
__device__ double kirkland_v(const double* p, double r) {
  const double pi = 3.141592653589793;
  const double pref = 4.0 * pi * pi * 0.529 * 14.4;
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double a = p[i], b = p[3 + i], c = p[6 + i], d = p[9 + i];
    v += pref * a * k0_integral(2.0 * pi * r * sqrt(b) + 0.1);
    v += pref / 2.0 * (c / d) * exp(-pi * pi * r * r / d);
  }
  return v;
}

// Mean of the os x os subsamples of table pixel i; subsample axes are
// linspace(-range/2, range/2, os*P) (the mgrid of msa.py:141-143).
__device__ double table_bin(const double* p, int i, int Py, int Px, int os, double range) {
  const int r = i / Px, c = i % Px;
  const double half = range / 2.0;
  const double s0 = range / double(os * Py - 1), s1 = range / double(os * Px - 1);
  double acc = 0.0;
  for (int dy = 0; dy < os; ++dy) {
    const double gy = -half + double(os * r + dy) * s0;
    double row = 0.0;
    for (int dx = 0; dx < os; ++dx) {
      const double gx = -half + double(os * c + dx) * s1;
      row += kirkland_v(p, sqrt(gy * gy + gx * gx));
    }
    acc += row / os;
  }
  return acc / os;
}

constexpr int kTableThreads = 256;

__global__ void __launch_bounds__(kTableThreads)
atomic_tables_kernel(const double* __restrict__ params, int Py, int Px, int os, double range,
                     float* __restrict__ tables) {
  __shared__ double red[kTableThreads / 64];
  const double* p = params + blockIdx.x * 12;
  const int n = Py * Px;
  double vmin = 1.0e300;
  for (int i = threadIdx.x; i < n; i += kTableThreads)
    vmin = fmin(vmin, table_bin(p, i, Py, Px, os, range));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vmin = fmin(vmin, __shfl_xor(vmin, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vmin;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < kTableThreads / 64; ++w) m = fmin(m, red[w]);
  float* out = tables + static_cast<long long>(blockIdx.x) * n;
  for (int i = threadIdx.x; i < n; i += kTableThreads)
    out[i] = float(table_bin(p, i, Py, Px, os, range) - m);
}

// ------------------------------------------------------------------ slice builder
constexpr int kTileH = 16;
constexpr int kTileW = 64;
constexpr int kSliceThreads = 256;
constexpr int kPxPerThread = kTileH * kTileW / kSliceThreads;  // 4

struct SiteBox {
  int z, y0, r_lo, r_hi, x0, c_lo, c_hi;
};

// POT: write the projected potential V (float) instead of t = exp(i sigma V).
template <bool POT>
__global__ void __launch_bounds__(kSliceThreads)
build_slices_kernel(const float* __restrict__ tables, int nZ, int Py, int Px,
                    const int* __restrict__ sites,
                    const int* __restrict__ offsets, int Ny, int Nx, float sigma,
                    float2* __restrict__ t_out, float* __restrict__ v_out) {
  __shared__ SiteBox list[kSliceThreads];
  __shared__ int wave_count[kSliceThreads / 64];

  const int s = blockIdx.z;
  const int ty0 = blockIdx.y * kTileH, tx0 = blockIdx.x * kTileW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = tx0 + lane;
  float acc[kPxPerThread];
#pragma unroll
  for (int k = 0; k < kPxPerThread; ++k) acc[k] = 0.0f;

  const int beg = offsets[s], end = offsets[s + 1];
  for (int base = beg; base < end; base += kSliceThreads) {
    const int i = base + threadIdx.x;
    bool hit = false;
    SiteBox bx{};
    if (i < end) {
      const int z = sites[3 * i], yc = sites[3 * i + 1], xc = sites[3 * i + 2];
      bx.z = z;
      bx.y0 = static_cast<int>(rintf(float(yc) - Py * 0.5f));
      const int y1 = static_cast<int>(rintf(float(yc) + Py * 0.5f));
      bx.x0 = static_cast<int>(rintf(float(xc) - Px * 0.5f));
      const int x1 = static_cast<int>(rintf(float(xc) + Px * 0.5f));
      bx.r_lo = max(bx.y0, 0);
      bx.r_hi = min(min(y1, Ny), bx.y0 + Py);
      bx.c_lo = max(bx.x0, 0);
      bx.c_hi = min(min(x1, Nx), bx.x0 + Px);
      hit = z >= 0 && z < nZ && bx.r_lo < bx.r_hi && bx.c_lo < bx.c_hi && bx.r_lo < ty0 + kTileH &&
            bx.r_hi > ty0 && bx.c_lo < tx0 + kTileW && bx.c_hi > tx0;
    }
    // order-preserving compaction: rank = hits in lower lanes + hits in lower waves
    const unsigned long long ballot = __ballot(hit);
    const int rank_in_wave = __builtin_amdgcn_mbcnt_hi(
        static_cast<unsigned>(ballot >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(ballot), 0));
    if (lane == 0) wave_count[wave] = __popcll(ballot);
    __syncthreads();
    int offset = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kSliceThreads / 64; ++w) {
      if (w < wave) offset += wave_count[w];
      total += wave_count[w];
    }
    if (hit) list[offset + rank_in_wave] = bx;
    __syncthreads();
    for (int e = 0; e < total; ++e) {
      const SiteBox b = list[e];
      if (col < b.c_lo || col >= b.c_hi) continue;
      const float* tab = tables + static_cast<long long>(b.z) * Py * Px + (col - b.x0);
#pragma unroll
      for (int k = 0; k < kPxPerThread; ++k) {
        const int row = ty0 + wave + 4 * k;
        if (row >= b.r_lo && row < b.r_hi) acc[k] += tab[(row - b.y0) * Px];
      }
    }
    __syncthreads();
  }

  const long long plane = static_cast<long long>(Ny) * Nx;
#pragma unroll
  for (int k = 0; k < kPxPerThread; ++k) {
    const int row = ty0 + wave + 4 * k;
    if (row < Ny && col < Nx) {
      if constexpr (POT) {
        v_out[s * plane + static_cast<long long>(row) * Nx + col] = acc[k];
        continue;
      }
      float sn, cs;
      sincosf(acc[k] * sigma, &sn, &cs);
      if (isnan(cs)) cs = 1.0f;
      if (isnan(sn)) sn = 1.0f;
      t_out[s * plane + static_cast<long long>(row) * Nx + col] = make_float2(cs, sn);
    }
  }
}

// ------------------------------------------------- sampled 3D potentials
// Slice s of a potential sampled at z[0..Z) integrates samples [s d, (s+1) d)
// with the trapezoid rule on their own coordinates and negates the result
// (MSA.integrate_potential_slices, namsa/msa.py:390-407: np.split into S
// chunks, trapz over each, V = -integral; no interval spans two chunks), and
// t = exp(i sigma V) (MSAGPU.integrate_potential_slices, msa.py:568-571).
// Each term is (z1 - z0) * (p0 + p1) / 2, summed in z order without
// contraction, as numpy's trapz evaluates it.
__global__ void __launch_bounds__(256) integrate_slices_kernel(const double* pot, const double* z, int d,
                                                               long long plane, double sigma, float2* t_out,
                                                               double* v_out) {
#pragma clang fp contract(off)
  const long long p = blockIdx.x * 256LL + threadIdx.x;
  if (p >= plane) return;
  const long long s = blockIdx.y;
  const double* c = pot + s * d * plane + p;
  const double* zs = z + s * d;
  double acc = 0.0;
  double prev = c[0];
  for (int j = 1; j < d; ++j) {
    const double cur = c[j * plane];
    acc += (zs[j] - zs[j - 1]) * (prev + cur) / 2.0;
    prev = cur;
  }
  const double v = -acc;
  const long long o = s * plane + p;
  if (v_out) v_out[o] = v;
  if (t_out) {
    double sn, cs;
    sincos(sigma * v, &sn, &cs);
    t_out[o] = make_float2(static_cast<float>(cs), static_cast<float>(sn));
  }
}

}  // namespace namsa

using namespace namsa;

extern "C" {

int namsa_integrate_slices(const double* pot, const double* z, int S, int d, int Ny, int Nx, double sigma,
                           void* t_out, double* v_out, void* stream) {
  if (!pot || !z || (!t_out && !v_out) || S < 0 || d < 1 || Ny <= 0 || Nx <= 0)
    return fail(kErrArg, "namsa_integrate_slices: bad arguments (S=%d d=%d Ny=%d Nx=%d)", S, d, Ny, Nx);
  if (S == 0) return kOk;
  if (S > 65535) return fail(kErrArg, "namsa_integrate_slices: S=%d exceeds the grid y limit", S);
  const long long plane = static_cast<long long>(Ny) * Nx;
  const long long blocks = (plane + 255) / 256;
  if (blocks > 0x7fffffffLL) return fail(kErrArg, "namsa_integrate_slices: plane too large");
  hipLaunchKernelGGL(integrate_slices_kernel, dim3(static_cast<unsigned>(blocks), S), dim3(256), 0,
                     as_stream(stream), pot, z, d, plane, sigma, static_cast<float2*>(t_out), v_out);
  return check_hip(hipGetLastError(), "integrate_slices_kernel launch");
}

}  // extern "C"

extern "C" {

int namsa_atomic_tables(const double* params, int nZ, int Py, int Px, double range, int oversample,
                        float* tables, void* stream) {
  if (!params || !tables || nZ <= 0 || Py <= 0 || Px <= 0 || !(range > 0) || oversample < 1 ||
      oversample * static_cast<long long>(Py) < 2 || oversample * static_cast<long long>(Px) < 2)
    return fail(kErrArg, "namsa_atomic_tables: bad arguments (nZ=%d Py=%d Px=%d os=%d)", nZ, Py,
                Px, oversample);
  hipLaunchKernelGGL(atomic_tables_kernel, dim3(nZ), dim3(kTableThreads), 0, as_stream(stream),
                     params, Py, Px, oversample, range, tables);
  return check_hip(hipGetLastError(), "atomic_tables_kernel launch");
}

int namsa_build_slices(const float* tables, int nZ, int Py, int Px, const int* sites,
                       const int* slice_offsets, int S, int Ny, int Nx, float sigma, void* t_out,
                       void* stream) {
  if (!tables || !slice_offsets || !t_out || nZ <= 0 || Py <= 0 || Px <= 0 || S < 0 || Ny <= 0 ||
      Nx <= 0)
    return fail(kErrArg, "namsa_build_slices: bad arguments");
  if (S == 0) return kOk;
  if (S > 65535) return fail(kErrArg, "namsa_build_slices: S=%d exceeds the grid z limit", S);
  dim3 grid((Nx + kTileW - 1) / kTileW, (Ny + kTileH - 1) / kTileH, S);
  hipLaunchKernelGGL(build_slices_kernel<false>, grid, dim3(kSliceThreads), 0, as_stream(stream), tables,
                     nZ, Py, Px, sites, slice_offsets, Ny, Nx, sigma, static_cast<float2*>(t_out), nullptr);
  return check_hip(hipGetLastError(), "build_slices_kernel launch");
}

int namsa_build_potential(const float* tables, int nZ, int Py, int Px, const int* sites,
                          const int* slice_offsets, int S, int Ny, int Nx, float* v_out, void* stream) {
  if (!tables || !slice_offsets || !v_out || nZ <= 0 || Py <= 0 || Px <= 0 || S < 0 || Ny <= 0 ||
      Nx <= 0)
    return fail(kErrArg, "namsa_build_potential: bad arguments");
  if (S == 0) return kOk;
  if (S > 65535) return fail(kErrArg, "namsa_build_potential: S=%d exceeds the grid z limit", S);
  dim3 grid((Nx + kTileW - 1) / kTileW, (Ny + kTileH - 1) / kTileH, S);
  hipLaunchKernelGGL(build_slices_kernel<true>, grid, dim3(kSliceThreads), 0, as_stream(stream), tables,
                     nZ, Py, Px, sites, slice_offsets, Ny, Nx, 0.0f, nullptr, v_out);
  return check_hip(hipGetLastError(), "build_slices_kernel launch");
}

}  // extern "C"
