// Fused batched multislice for gfx950: probe former, slice step, readout.
//
// One slice of the reference (namsa/msa.py:971-981) is
//     psi <- IFFT_u( PM * FFT( psi * t_i / N^2 ) ).
// With the 2D transform split into a row transform R (along x) and a column
// transform C (along y), FFT = C R and IFFT_u = R^-1 C^-1, and the chain of
// slices regroups into two passes over the wave per slice:
//     row pass    u = R( t_i * R^-1(v) / N^2 )        (R^-1 of the previous slice fused in)
//     column pass v = C^-1( PM * C(u) )
// Each pass reads and writes the complex64 wave once, 32*N^2 bytes per
// probe-slice in total, against ~112*N^2 for the reference's four launches.
//
// The probe stack fftshift(IFFT_u(psi_k * ramp)) (msa.py:959-964) enters as
// v0 = C^-1(psi_k * ramp * (-1)^(ky+kx)) -- for even sizes the real-space
// fftshift is the (-1)^k modulation of the spectrum -- so the first row pass
// picks it up like any other slice.  The readout |sqrt(N^2) FFT(psi/N^2)|^2
// (msa.py:985-990) equals N^2 |PM * C(u_last)|^2 because FFT IFFT_u = N^2,
// so the last column pass writes intensities instead of transforming back.
#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "fft.hpp"
#include "namsa_hip.h"

namespace namsa {

// ------------------------------------------------------------------ schedules
// Row transforms: T = N/E threads per row; T <= 64 keeps a row inside one
// wavefront so its LDS exchanges need no workgroup barrier.
template <int N>
struct RowCfg;
template <> struct RowCfg<128>  { static constexpr int E = 8;  using RL = Radices<8, 4, 4>; };
template <> struct RowCfg<192>  { static constexpr int E = 12; using RL = Radices<3, 4, 4, 4>; };
template <> struct RowCfg<256>  { static constexpr int E = 8;  using RL = Radices<8, 8, 4>; };
template <> struct RowCfg<384>  { static constexpr int E = 12; using RL = Radices<3, 2, 4, 4, 4>; };
template <> struct RowCfg<512>  { static constexpr int E = 8;  using RL = Radices<8, 8, 8>; };
template <> struct RowCfg<768>  { static constexpr int E = 12; using RL = Radices<3, 4, 4, 4, 4>; };
template <> struct RowCfg<1024> { static constexpr int E = 16; using RL = Radices<16, 8, 8>; };
template <> struct RowCfg<2048> { static constexpr int E = 16; using RL = Radices<16, 16, 8>; };

// Column transforms: C adjacent columns per workgroup so that each global
// access of a wavefront covers C*8 contiguous bytes of C*8/64 rows.
template <int N>
struct ColCfg;
template <> struct ColCfg<128>  { static constexpr int E = 8;  static constexpr int C = 16; using RL = Radices<8, 4, 4>; };
template <> struct ColCfg<192>  { static constexpr int E = 12; static constexpr int C = 16; using RL = Radices<3, 4, 4, 4>; };
template <> struct ColCfg<256>  { static constexpr int E = 8;  static constexpr int C = 16; using RL = Radices<8, 8, 4>; };
template <> struct ColCfg<384>  { static constexpr int E = 12; static constexpr int C = 16; using RL = Radices<3, 2, 4, 4, 4>; };
template <> struct ColCfg<512>  { static constexpr int E = 16; static constexpr int C = 16; using RL = Radices<8, 8, 8>; };
template <> struct ColCfg<768>  { static constexpr int E = 12; static constexpr int C = 8;  using RL = Radices<3, 4, 4, 4, 4>; };
template <> struct ColCfg<1024> { static constexpr int E = 16; static constexpr int C = 8;  using RL = Radices<16, 8, 8>; };
template <> struct ColCfg<2048> { static constexpr int E = 16; static constexpr int C = 8;  using RL = Radices<16, 16, 8>; };

constexpr int kRowThreads = 256;

template <int N>
constexpr int col_threads() { return ColCfg<N>::C * (N / ColCfg<N>::E); }

enum RowOp { ROW_STEP = 0, ROW_INV = 1, ROW_FWD = 2 };
enum ColOp { COL_PROP = 0, COL_PROBE = 1, COL_READOUT = 2, COL_FWD = 3, COL_INV = 4 };

// ------------------------------------------------------------------ row pass
// psi: [n_rows][N] (rows of all images back to back); tsl: [rows_per_image][N].
template <int N, int OP>
__global__ void __launch_bounds__(kRowThreads)
row_kernel(float2* __restrict__ psi, long long n_rows, int rows_per_image,
           const float2* __restrict__ tsl, float scale) {
  constexpr int E = RowCfg<N>::E;
  using F = Fft<N, E, typename RowCfg<N>::RL>;
  constexpr int T = F::T;
  constexpr int G = kRowThreads / T;
  using Sync = typename std::conditional<(T <= 64), WaveSync, BlockSync>::type;
  __shared__ float2 smem[G * RowLds::padded(N)];

  const int g = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const RowLds lds{smem + g * RowLds::padded(N)};
  float2 tw[F::NTW_ALLOC];
  F::twiddles(t, tw);

  const long long n_blocks = (n_rows + G - 1) / G;
  for (long long rb = blockIdx.x; rb < n_blocks; rb += gridDim.x) {
    const long long row = rb * G + g;
    const bool active = row < n_rows;
    const long long rr = active ? row : n_rows - 1;
    float2* p = psi + rr * N;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = p[t + T * m];
    if constexpr (OP == ROW_STEP) {
      const float2* tr = tsl + static_cast<long long>(rr % rows_per_image) * N;
      float2 tv[E];
#pragma unroll
      for (int m = 0; m < E; ++m) tv[m] = tr[t + T * m];
      F::template run<true, Sync>(v, lds, t, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cscale(cmul(v[m], tv[m]), scale);
      F::template run<false, Sync>(v, lds, t, tw);
    } else if constexpr (OP == ROW_INV) {
      F::template run<true, Sync>(v, lds, t, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cscale(v[m], scale);
    } else {
      F::template run<false, Sync>(v, lds, t, tw);
    }
    if (active) {
#pragma unroll
      for (int m = 0; m < E; ++m) p[t + T * m] = v[m];
    }
  }
}

// ------------------------------------------------------------------ column pass
struct ColArgs {
  float2* psi;          // [B][Ny][Nx]
  int Nx;
  int tiles_per_image;  // Nx / C
  long long n_tiles;    // B * tiles_per_image
  const float2* pm;     // [Ny][Nx] propagator * mask (PROP, READOUT)
  const float2* psi_k;  // [Ny][Nx] (PROBE)
  const float* pos;     // [B][2] (y, x) (PROBE)
  double kmax;          // (PROBE)
  float amp;            // intensity factor (READOUT)
  int apply_pm;         // (READOUT)
  float* cbed;          // [B][Ny][Nx] (READOUT)
  const float* det;     // [n_det][Ny][Nx] (READOUT)
  int n_det;
  double* det_ws;       // [n_tiles][n_det] (READOUT)
};

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// exp(2 pi i * turns) with the phase reduced to [-1/2, 1/2] turns in double.
__device__ __forceinline__ float2 unit_phase(double turns) {
  const double f = turns - rint(turns);
  double s, c;
  sincospi(2.0 * f, &s, &c);
  return make_float2(float(c), float(s));
}

template <int N, int OP>
__global__ void __launch_bounds__(col_threads<N>())
col_kernel(ColArgs a) {
  constexpr int E = ColCfg<N>::E;
  constexpr int C = ColCfg<N>::C;
  using F = Fft<N, E, typename ColCfg<N>::RL>;
  constexpr int T = F::T;
  constexpr int NT = C * T;
  constexpr int NW = NT / 64;
  __shared__ float2 smem[ColLds<C>::padded(N)];
  __shared__ double red[NW];

  const int c = threadIdx.x % C;
  const int jt = threadIdx.x / C;
  const ColLds<C> lds{smem, c};
  float2 tw[F::NTW_ALLOC];
  F::twiddles(jt, tw);
  const long long plane = static_cast<long long>(N) * a.Nx;

  for (long long tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
    const long long b = tile / a.tiles_per_image;
    const int col = static_cast<int>(tile % a.tiles_per_image) * C + c;
    float2* base = a.psi + b * plane + col;
    float2 v[E];

    if constexpr (OP == COL_PROBE) {
      // psi_k * exp(2 pi i (-kx rx - ky ry)) * (-1)^(row + col)
      // (build_probes_stack + fftshift_2d_stack, probe_kernels.cu:239-304)
      const double ry = a.pos[2 * b];
      const double rx = a.pos[2 * b + 1];
      const float2 ex = unit_phase(-k_lin(col, a.kmax, a.Nx) * rx);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int row = jt + T * m;
        float2 e = cmul(ex, unit_phase(-k_lin(row, a.kmax, N) * ry));
        if ((row + col) & 1) e = make_float2(-e.x, -e.y);
        v[m] = cmul(a.psi_k[static_cast<long long>(row) * a.Nx + col], e);
      }
      F::template run<true, BlockSync>(v, lds, jt, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) base[static_cast<long long>(jt + T * m) * a.Nx] = v[m];
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = base[static_cast<long long>(jt + T * m) * a.Nx];

      if constexpr (OP == COL_PROP) {
        F::template run<false, BlockSync>(v, lds, jt, tw);
#pragma unroll
        for (int m = 0; m < E; ++m)
          v[m] = cmul(v[m], a.pm[static_cast<long long>(jt + T * m) * a.Nx + col]);
        F::template run<true, BlockSync>(v, lds, jt, tw);
      } else if constexpr (OP == COL_FWD) {
        F::template run<false, BlockSync>(v, lds, jt, tw);
      } else if constexpr (OP == COL_INV) {
        F::template run<true, BlockSync>(v, lds, jt, tw);
      }

      if constexpr (OP == COL_READOUT) {
        F::template run<false, BlockSync>(v, lds, jt, tw);
        float inten[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const long long off = static_cast<long long>(jt + T * m) * a.Nx + col;
          float2 y = v[m];
          if (a.apply_pm) y = cmul(y, a.pm[off]);
          inten[m] = (y.x * y.x + y.y * y.y) * a.amp;
          a.cbed[b * plane + off] = inten[m];
        }
        for (int d = 0; d < a.n_det; ++d) {
          const float* w = a.det + d * plane;
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < E; ++m)
            s += double(inten[m]) * double(w[static_cast<long long>(jt + T * m) * a.Nx + col]);
          s = wave_sum(s);
          if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
          __syncthreads();
          if (threadIdx.x == 0) {
            double tot = 0.0;
#pragma unroll
            for (int i = 0; i < NW; ++i) tot += red[i];
            a.det_ws[tile * a.n_det + d] = tot;
          }
          __syncthreads();
        }
      } else {
#pragma unroll
        for (int m = 0; m < E; ++m) base[static_cast<long long>(jt + T * m) * a.Nx] = v[m];
      }
    }
  }
}

// det_out[b][d] = sum over the column tiles of image b, in tile order.
__global__ void det_finalize_kernel(const double* __restrict__ ws, int tiles_per_image, int n_det,
                                    int B, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * n_det) return;
  const int b = i / n_det, d = i % n_det;
  double s = 0.0;
  for (int k = 0; k < tiles_per_image; ++k)
    s += ws[(static_cast<long long>(b) * tiles_per_image + k) * n_det + d];
  out[i] = s;
}

// ------------------------------------------------------------------ small kernels
// psi_k and aperture (probe_kernels.cu:180-222): k in double on the linspace
// grid, stored as float; chi in double with the reference's float constants.
__global__ void probe_k_kernel(int Ny, int Nx, double kmax, float k_semi, float lam, float C1,
                               float C3, float C5, int smooth, float steep,
                               float2* __restrict__ psi_k, float* __restrict__ apert) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Ny * Nx) return;
  const int row = i / Nx, col = i % Nx;
  const double kx = k_lin(col, kmax, Nx), ky = k_lin(row, kmax, Ny);
  const double k = sqrt(kx * kx + ky * ky);
  const float kf = float(k);
  double ap;
  if (smooth)
    ap = 1.0 / (1.0 + exp(-2.0 * double(steep) * (double(k_semi) - double(kf))));
  else
    ap = (kf < k_semi) ? 1.0 : 0.0;
  const double pi_f = double(3.14159265358979f);  // acos(-1.f), probe_kernels.cu:210
  const double c1 = double(-0.5f * C1), c3 = double(0.25f * C3), c5 = double((1.f / 6) * C5);
  const double kl = k * double(lam);
  const double kl2 = kl * kl;
  const double chi = 2.0 * pi_f / double(lam) * (c1 * kl2 + c3 * kl2 * kl2 + c5 * kl2 * kl2 * kl2);
  double s, c;
  sincos(chi, &s, &c);
  psi_k[i] = make_float2(float(c * ap), float(-s * ap));
  if (apert) apert[i] = float(ap);
}

// PM = exp(-i pi k^2 lambda dz) * [k < k_cut] (probe_kernels.cu:167-194).
__global__ void propagator_kernel(int Ny, int Nx, double kmax, float k_cut, float lam, float dz,
                                  float2* __restrict__ pm, float* __restrict__ mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Ny * Nx) return;
  const int row = i / Nx, col = i % Nx;
  const double kx = k_lin(col, kmax, Nx), ky = k_lin(row, kmax, Ny);
  const float kf = float(sqrt(kx * kx + ky * ky));
  const float msk = (kf < k_cut) ? 1.0f : 0.0f;
  const double ph = 3.141592653589793 * double(kf) * double(kf) * double(lam) * double(dz);
  double s, c;
  sincos(ph, &s, &c);
  pm[i] = make_float2(float(c) * msk, float(-s) * msk);
  if (mask) mask[i] = msk;
}

// out[b][d] = sum_pixels cbed[b] * det[d]; one workgroup per (b, d).
__global__ void __launch_bounds__(256)
integrate_kernel(const float* __restrict__ cbed, long long plane, const float* __restrict__ det,
                 int n_det, double* __restrict__ out) {
  __shared__ double red[4];
  const long long b = blockIdx.x;
  const int d = blockIdx.y;
  const float* img = cbed + b * plane;
  const float* w = det + d * plane;
  double s = 0.0;
  for (long long i = threadIdx.x; i < plane; i += 256) s += double(img[i]) * double(w[i]);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[b * n_det + d] = red[0] + red[1] + red[2] + red[3];
}

__global__ void norm2_kernel(const float2* __restrict__ x, long long n, double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    const float2 v = x[i];
    s += double(v.x) * double(v.x) + double(v.y) * double(v.y);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) tot += red[i];
    *out = tot;
  }
}

// ------------------------------------------------------------------ launchers
template <class K>
int persistent_grid(K kernel, int threads, long long work) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  const long long cap = static_cast<long long>(per_cu) * device_cu_count();
  return static_cast<int>(std::max<long long>(1, std::min(work, cap)));
}

template <int N, int OP>
int launch_row(float2* psi, long long n_rows, int rows_per_image, const float2* tsl, float scale,
               hipStream_t st) {
  auto k = row_kernel<N, OP>;
  constexpr int G = kRowThreads / (N / RowCfg<N>::E);
  static const int cap = persistent_grid(k, kRowThreads, 1LL << 40);
  const long long blocks = (n_rows + G - 1) / G;
  const int grid = static_cast<int>(std::min<long long>(blocks, cap));
  hipLaunchKernelGGL(k, dim3(grid), dim3(kRowThreads), 0, st, psi, n_rows, rows_per_image, tsl,
                     scale);
  return check_hip(hipGetLastError(), "row_kernel launch");
}

template <int N, int OP>
int launch_col(const ColArgs& a, hipStream_t st) {
  auto k = col_kernel<N, OP>;
  static const int cap = persistent_grid(k, col_threads<N>(), 1LL << 40);
  const int grid = static_cast<int>(std::min<long long>(a.n_tiles, cap));
  hipLaunchKernelGGL(k, dim3(grid), dim3(col_threads<N>()), 0, st, a);
  return check_hip(hipGetLastError(), "col_kernel launch");
}

#define NAMSA_FOR_EACH_N(X) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(2048)

bool supported(int n) {
  switch (n) {
#define X(n) case n:
    NAMSA_FOR_EACH_N(X)
#undef X
    return true;
    default:
      return false;
  }
}

int row_pass(int Nx, int op, float2* psi, long long n_rows, int rows_per_image,
             const float2* tsl, float scale, hipStream_t st) {
  switch (Nx) {
#define X(n)                                                                               \
  case n:                                                                                  \
    if (op == ROW_STEP) return launch_row<n, ROW_STEP>(psi, n_rows, rows_per_image, tsl, scale, st); \
    if (op == ROW_INV) return launch_row<n, ROW_INV>(psi, n_rows, rows_per_image, tsl, scale, st);   \
    return launch_row<n, ROW_FWD>(psi, n_rows, rows_per_image, tsl, scale, st);
    NAMSA_FOR_EACH_N(X)
#undef X
    default:
      return fail(kErrUnsupported, "unsupported row length %d", Nx);
  }
}

int col_pass(int Ny, int op, const ColArgs& a, hipStream_t st) {
  switch (Ny) {
#define X(n)                                                       \
  case n:                                                          \
    switch (op) {                                                  \
      case COL_PROP: return launch_col<n, COL_PROP>(a, st);        \
      case COL_PROBE: return launch_col<n, COL_PROBE>(a, st);      \
      case COL_READOUT: return launch_col<n, COL_READOUT>(a, st);  \
      case COL_FWD: return launch_col<n, COL_FWD>(a, st);          \
      default: return launch_col<n, COL_INV>(a, st);               \
    }
    NAMSA_FOR_EACH_N(X)
#undef X
    default:
      return fail(kErrUnsupported, "unsupported column length %d", Ny);
  }
}

int col_tile(int Ny) {
  switch (Ny) {
#define X(n) case n: return ColCfg<n>::C;
    NAMSA_FOR_EACH_N(X)
#undef X
    default:
      return 0;
  }
}

ColArgs col_args(float2* psi, int B, int Ny, int Nx) {
  ColArgs a{};
  a.psi = psi;
  a.Nx = Nx;
  a.tiles_per_image = Nx / col_tile(Ny);
  a.n_tiles = static_cast<long long>(B) * a.tiles_per_image;
  return a;
}

int check_grid(int Ny, int Nx, const char* who) {
  if (!supported(Ny) || !supported(Nx))
    return fail(kErrUnsupported, "%s: grid %dx%d is not a compiled transform size", who, Ny, Nx);
  return kOk;
}

}  // namespace namsa

using namespace namsa;

extern "C" {

int namsa_fft_supported(int n) { return supported(n) ? 1 : 0; }

int namsa_fft2d(void* data, int B, int Ny, int Nx, int inverse, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_fft2d")) return rc;
  if (!data || B <= 0) return fail(kErrArg, "namsa_fft2d: null data or B=%d", B);
  hipStream_t st = as_stream(stream);
  float2* p = static_cast<float2*>(data);
  ColArgs a = col_args(p, B, Ny, Nx);
  const long long rows = static_cast<long long>(B) * Ny;
  if (!inverse) {
    if (int rc = row_pass(Nx, ROW_FWD, p, rows, Ny, nullptr, 1.0f, st)) return rc;
    return col_pass(Ny, COL_FWD, a, st);
  }
  if (int rc = col_pass(Ny, COL_INV, a, st)) return rc;
  return row_pass(Nx, ROW_INV, p, rows, Ny, nullptr, 1.0f, st);
}

int namsa_probe_k(int Ny, int Nx, float kmax, float k_semi, float lambda, float C1, float C3,
                  float C5, int smooth, float steepness, void* psi_k, float* apert,
                  void* stream) {
  if (Ny < 2 || Nx < 2 || !psi_k) return fail(kErrArg, "namsa_probe_k: bad arguments");
  const int n = Ny * Nx;
  hipLaunchKernelGGL(probe_k_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), Ny,
                     Nx, double(kmax), k_semi, lambda, C1, C3, C5, smooth, steepness,
                     static_cast<float2*>(psi_k), apert);
  return check_hip(hipGetLastError(), "probe_k_kernel launch");
}

int namsa_propagator(int Ny, int Nx, float kmax, float k_cut, float lambda, float dz, void* pm,
                     float* mask, void* stream) {
  if (Ny < 2 || Nx < 2 || !pm) return fail(kErrArg, "namsa_propagator: bad arguments");
  const int n = Ny * Nx;
  hipLaunchKernelGGL(propagator_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream),
                     Ny, Nx, double(kmax), k_cut, lambda, dz, static_cast<float2*>(pm), mask);
  return check_hip(hipGetLastError(), "propagator_kernel launch");
}

int namsa_norm2(const void* data, long long n, double* out, void* stream) {
  if (!data || !out || n <= 0) return fail(kErrArg, "namsa_norm2: bad arguments");
  hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, as_stream(stream),
                     static_cast<const float2*>(data), n, out);
  return check_hip(hipGetLastError(), "norm2_kernel launch");
}

int namsa_integrate(const float* cbed, int B, int Ny, int Nx, const float* det, int nDet,
                    double* out, void* stream) {
  if (!cbed || !det || !out || B <= 0 || Ny <= 0 || Nx <= 0 || nDet <= 0 || nDet > 65535)
    return fail(kErrArg, "namsa_integrate: bad arguments");
  hipLaunchKernelGGL(integrate_kernel, dim3(B, nDet), dim3(256), 0, as_stream(stream), cbed,
                     static_cast<long long>(Ny) * Nx, det, nDet, out);
  return check_hip(hipGetLastError(), "integrate_kernel launch");
}

size_t namsa_det_workspace_doubles(int B, int Ny, int Nx, int nDet) {
  const int c = col_tile(Ny);
  if (c == 0 || B <= 0 || nDet <= 0) return 0;
  return static_cast<size_t>(B) * (Nx / c) * nDet;
}

int namsa_probe_init(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                     void* psi_ws, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_probe_init")) return rc;
  if (B <= 0 || !psi_k || !pos_yx || !psi_ws) return fail(kErrArg, "namsa_probe_init: bad arguments");
  ColArgs a = col_args(static_cast<float2*>(psi_ws), B, Ny, Nx);
  a.psi_k = static_cast<const float2*>(psi_k);
  a.pos = pos_yx;
  a.kmax = double(kmax);
  return col_pass(Ny, COL_PROBE, a, as_stream(stream));
}

int namsa_slice_rows(void* psi_ws, int B, int Ny, int Nx, const void* t_slice, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_slice_rows")) return rc;
  if (B <= 0 || !psi_ws || !t_slice) return fail(kErrArg, "namsa_slice_rows: bad arguments");
  const float inv_n2 = 1.0f / static_cast<float>(static_cast<long long>(Ny) * Nx);
  return row_pass(Nx, ROW_STEP, static_cast<float2*>(psi_ws), static_cast<long long>(B) * Ny, Ny,
                  static_cast<const float2*>(t_slice), inv_n2, as_stream(stream));
}

int namsa_slice_cols(void* psi_ws, int B, int Ny, int Nx, const void* pm, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_slice_cols")) return rc;
  if (B <= 0 || !psi_ws || !pm) return fail(kErrArg, "namsa_slice_cols: bad arguments");
  ColArgs a = col_args(static_cast<float2*>(psi_ws), B, Ny, Nx);
  a.pm = static_cast<const float2*>(pm);
  return col_pass(Ny, COL_PROP, a, as_stream(stream));
}

int namsa_multislice(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                     const void* t, int S, const void* pm, int mode, float out_scale,
                     void* psi_ws, float* cbed, const float* det, int nDet, double* det_ws,
                     double* det_out, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_multislice")) return rc;
  if (B <= 0 || S < 0 || !psi_k || !pos_yx || !psi_ws)
    return fail(kErrArg, "namsa_multislice: B=%d S=%d or null probe/positions/workspace", B, S);
  if (mode != NAMSA_OUT_CBED && mode != NAMSA_OUT_INCIDENT && mode != NAMSA_OUT_EXIT)
    return fail(kErrArg, "namsa_multislice: unknown mode %d", mode);
  if (mode != NAMSA_OUT_INCIDENT && S > 0 && (!t || !pm))
    return fail(kErrArg, "namsa_multislice: null slices or propagator");
  if (mode == NAMSA_OUT_CBED && !cbed) return fail(kErrArg, "namsa_multislice: null cbed");
  if (mode == NAMSA_OUT_CBED && nDet > 0 && (!det || !det_ws || !det_out))
    return fail(kErrArg, "namsa_multislice: null detector buffers");

  hipStream_t st = as_stream(stream);
  float2* psi = static_cast<float2*>(psi_ws);
  const float2* tt = static_cast<const float2*>(t);
  const long long plane = static_cast<long long>(Ny) * Nx;
  const long long rows = static_cast<long long>(B) * Ny;
  const float inv_n2 = 1.0f / static_cast<float>(plane);

  ColArgs a = col_args(psi, B, Ny, Nx);
  a.psi_k = static_cast<const float2*>(psi_k);
  a.pos = pos_yx;
  a.kmax = double(kmax);
  a.pm = static_cast<const float2*>(pm);
  if (int rc = col_pass(Ny, COL_PROBE, a, st)) return rc;

  if (mode == NAMSA_OUT_INCIDENT) return row_pass(Nx, ROW_INV, psi, rows, Ny, nullptr, out_scale, st);

  const int last_prop = (mode == NAMSA_OUT_EXIT) ? S : S - 1;
  for (int i = 0; i < S; ++i) {
    if (int rc = row_pass(Nx, ROW_STEP, psi, rows, Ny, tt + i * plane, inv_n2, st)) return rc;
    if (i < last_prop)
      if (int rc = col_pass(Ny, COL_PROP, a, st)) return rc;
  }
  if (mode == NAMSA_OUT_EXIT) return row_pass(Nx, ROW_INV, psi, rows, Ny, nullptr, out_scale, st);

  // CBED readout.  S > 0: N^2 |PM C(u)|^2 / norm.  S == 0: the state is v0 and
  // FFT(psi0) = Nx C(v0), so |sqrt(N^2) FFT(psi0/N^2)|^2 = (Nx/Ny) |C(v0)|^2.
  a.apply_pm = S > 0 ? 1 : 0;
  a.amp = S > 0 ? static_cast<float>(static_cast<double>(plane) * out_scale)
                : static_cast<float>(static_cast<double>(Nx) / Ny * out_scale);
  a.cbed = cbed;
  a.det = det;
  a.n_det = nDet > 0 ? nDet : 0;
  a.det_ws = det_ws;
  if (int rc = col_pass(Ny, COL_READOUT, a, st)) return rc;
  if (nDet > 0) {
    const int n = B * nDet;
    hipLaunchKernelGGL(det_finalize_kernel, dim3((n + 255) / 256), dim3(256), 0, st, det_ws,
                       a.tiles_per_image, nDet, B, det_out);
    return check_hip(hipGetLastError(), "det_finalize_kernel launch");
  }
  return kOk;
}

}  // extern "C"
