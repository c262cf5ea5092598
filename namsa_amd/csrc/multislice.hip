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
#include <mutex>
#include <vector>

#include "common.hpp"
#include "namsa_hip.h"
#include "passes.hpp"

namespace namsa {

// Fixed device memory of namsa_multislice_noalloc: the wave state of one
// probe group followed by its detector partial sums.  Every GPU the library
// runs on holds its own copy until the process ends.  In host code the name
// is only a stand-in for the symbol; kernels get the per-device address that
// scratch_setup() looks up.
static __device__ __align__(256) unsigned char scratch[NAMSA_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // indexed by device ordinal

int scratch_setup() {
  int dev = 0;
  if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev >= static_cast<int>(g_scratch_ptr.size())) g_scratch_ptr.resize(dev + 1, nullptr);
  if (g_scratch_ptr[dev]) return kOk;
  void* p = nullptr;
  if (int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
    return rc;
  if (!p) return fail(kErrHip, "hipGetSymbolAddress(scratch) returned a null address");
  g_scratch_ptr[dev] = p;
  return kOk;
}

// Address of the current device's scratch array (looked up on first use).
static int scratch_address(unsigned char** out) {
  if (int rc = scratch_setup()) return rc;
  int dev = 0;
  if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return kOk;
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

// dst[tile_offset(y, x)] = src[y][x], one element per thread (reads coalesced).
__global__ void tile_plane_kernel(const float2* __restrict__ src, float2* __restrict__ dst, int Nx, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[tile_offset(i / Nx, i % Nx, Nx)] = src[i];
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

ColArgs col_args(float2* psi, int B, int Nx) {
  ColArgs a{};
  a.psi = psi;
  a.B = B;
  a.Nx = Nx;
  a.band_lo = 0;
  a.band_hi = Nx;
  a.pm_scale = 1.0f;
  return a;
}

RowArgs row_args(float2* psi, int B, int Ny, int Nx, const float2* tsl, float scale, int lo, int hi) {
  RowArgs r{};
  r.psi = psi;
  r.n_rows = static_cast<long long>(B) * Ny;
  r.rows_per_image = Ny;
  r.tsl = tsl;
  r.scale = scale;
  r.band_lo = lo;
  r.band_hi = hi;
  r.store_lo = 0;
  r.store_hi = Nx;
  return r;
}

int check_band(int Nx, int& lo, int& hi, const char* who) {
  if (lo == 0 && hi == 0) hi = Nx;
  if (lo < 0 || hi > Nx || lo >= hi)
    return fail(kErrArg, "%s: band [%d, %d) outside [0, %d)", who, lo, hi, Nx);
  return kOk;
}

// Thin-object exit wave probes * exp(i sum_s V_s) (namsa/utils.py:130-135),
// the phase summed over slices in slice order in FP64.
__global__ void __launch_bounds__(256) thin_object_kernel(const double* V, int S, const float2* probes, int B,
                                                          long long plane, float2* psi) {
  const long long p = blockIdx.x * 256LL + threadIdx.x;
  if (p >= plane) return;
  double ph = 0.0;
  for (int s = 0; s < S; ++s) ph += V[s * plane + p];
  double sn, cs;
  sincos(ph, &sn, &cs);
  for (int b = 0; b < B; ++b) {
    const float2 q = probes[b * plane + p];
    psi[b * plane + p] = make_float2(static_cast<float>(q.x * cs - q.y * sn), static_cast<float>(q.x * sn + q.y * cs));
  }
}

__global__ void __launch_bounds__(256) abs_kernel(const float2* psi, long long n, float* out) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += static_cast<long long>(gridDim.x) * 256) {
    const float2 q = psi[i];
    out[i] = sqrtf(q.x * q.x + q.y * q.y);
  }
}

int check_grid(int Ny, int Nx, const char* who) {
  if (!supported(Ny) || !supported(Nx))
    return fail(kErrUnsupported, "%s: grid %dx%d is not a compiled transform size", who, Ny, Nx);
  return kOk;
}

}  // namespace namsa

using namespace namsa;

extern "C" {

int namsa_fft2d(void* data, int B, int Ny, int Nx, int inverse, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_fft2d")) return rc;
  if (!data || B <= 0) return fail(kErrArg, "namsa_fft2d: null data or B=%d", B);
  hipStream_t st = as_stream(stream);
  float2* p = static_cast<float2*>(data);
  ColArgs a = col_args(p, B, Nx);
  const RowArgs r = row_args(p, B, Ny, Nx, nullptr, 1.0f, 0, Nx);
  if (!inverse) {
    if (int rc = row_pass(Nx, ROW_FWD, r, st)) return rc;
    return col_pass(Ny, COL_FWD, a, st);
  }
  if (int rc = col_pass(Ny, COL_INV, a, st)) return rc;
  return row_pass(Nx, ROW_INV, r, st);
}

int namsa_thin_object(const double* V, int S, const void* probes, int B, int Ny, int Nx, void* psi_ws, float* out,
                      void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_thin_object")) return rc;
  if (!V || !probes || !psi_ws || !out || S < 0 || B <= 0)
    return fail(kErrArg, "namsa_thin_object: bad arguments (S=%d B=%d)", S, B);
  hipStream_t st = as_stream(stream);
  const long long plane = static_cast<long long>(Ny) * Nx;
  hipLaunchKernelGGL(thin_object_kernel, dim3(static_cast<unsigned>((plane + 255) / 256)), dim3(256), 0, st, V, S,
                     static_cast<const float2*>(probes), B, plane, static_cast<float2*>(psi_ws));
  if (int rc = check_hip(hipGetLastError(), "thin_object_kernel launch")) return rc;
  if (int rc = namsa_fft2d(psi_ws, B, Ny, Nx, 0, stream)) return rc;
  const long long n = plane * B;
  const long long blocks = std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(abs_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                     static_cast<const float2*>(psi_ws), n, out);
  return check_hip(hipGetLastError(), "abs_kernel launch");
}

int namsa_row_fft(void* data, int B, int Ny, int Nx, int inverse, float scale, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_row_fft")) return rc;
  if (!data || B <= 0) return fail(kErrArg, "namsa_row_fft: null data or B=%d", B);
  if (!inverse && scale != 1.0f) return fail(kErrArg, "namsa_row_fft: the forward transform is unscaled");
  const RowArgs r = row_args(static_cast<float2*>(data), B, Ny, Nx, nullptr, scale, 0, Nx);
  return row_pass(Nx, inverse ? ROW_INV : ROW_FWD, r, as_stream(stream));
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
  if (!supported(Ny) || !supported(Nx) || B <= 0 || nDet <= 0) return 0;
  return static_cast<size_t>(B) * (Nx / kMinColTile) * nDet;
}

int namsa_probe_init(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                     void* psi_ws, void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_probe_init")) return rc;
  if (B <= 0 || !psi_k || !pos_yx || !psi_ws) return fail(kErrArg, "namsa_probe_init: bad arguments");
  ColArgs a = col_args(static_cast<float2*>(psi_ws), B, Nx);
  a.psi_k = static_cast<const float2*>(psi_k);
  a.pos = pos_yx;
  a.kmax = double(kmax);
  return col_pass(Ny, COL_PROBE, a, as_stream(stream));
}

}  // extern "C"

namespace {

int slice_rows(void* psi_ws, int B, int Ny, int Nx, const void* t_slice, int band_lo, int band_hi, int tiled,
               void* stream, const char* who) {
  if (int rc = check_grid(Ny, Nx, who)) return rc;
  if (B <= 0 || !psi_ws || !t_slice) return fail(kErrArg, "%s: bad arguments", who);
  if (int rc = check_band(Nx, band_lo, band_hi, who)) return rc;
  if (tiled && !(tiled_layout(Ny) && tiled_layout(Nx)))
    return fail(kErrUnsupported, "%s: no tiled layout at %d x %d", who, Ny, Nx);
  RowArgs r = row_args(static_cast<float2*>(psi_ws), B, Ny, Nx, static_cast<const float2*>(t_slice), 1.0f,
                       band_lo, band_hi);
  r.store_lo = band_lo / kStoreAlign * kStoreAlign;
  r.store_hi = std::min(Nx, (band_hi + kStoreAlign - 1) / kStoreAlign * kStoreAlign);
  r.tiled = tiled;
  return row_pass(Nx, ROW_STEP, r, as_stream(stream));
}

int slice_cols(void* psi_ws, int B, int Ny, int Nx, const void* pm, int band_lo, int band_hi, int tiled,
               void* stream, const char* who) {
  if (int rc = check_grid(Ny, Nx, who)) return rc;
  if (B <= 0 || !psi_ws || !pm) return fail(kErrArg, "%s: bad arguments", who);
  if (int rc = check_band(Nx, band_lo, band_hi, who)) return rc;
  if (tiled && !(tiled_layout(Ny) && tiled_layout(Nx)))
    return fail(kErrUnsupported, "%s: no tiled layout at %d x %d", who, Ny, Nx);
  ColArgs a = col_args(static_cast<float2*>(psi_ws), B, Nx);
  a.pm = static_cast<const float2*>(pm);
  a.pm_scale = 1.0f / static_cast<float>(static_cast<long long>(Ny) * Nx);
  a.band_lo = band_lo;
  a.band_hi = band_hi;
  a.tiled = tiled;
  a.pm_tiled = tiled;
  return col_pass(Ny, COL_PROP, a, as_stream(stream));
}

}  // namespace

extern "C" {

int namsa_slice_rows(void* psi_ws, int B, int Ny, int Nx, const void* t_slice, int band_lo, int band_hi,
                     void* stream) {
  return slice_rows(psi_ws, B, Ny, Nx, t_slice, band_lo, band_hi, 0, stream, "namsa_slice_rows");
}

int namsa_slice_cols(void* psi_ws, int B, int Ny, int Nx, const void* pm, int band_lo, int band_hi,
                     void* stream) {
  return slice_cols(psi_ws, B, Ny, Nx, pm, band_lo, band_hi, 0, stream, "namsa_slice_cols");
}

int namsa_wave_tiled(int Ny, int Nx) { return tiled_layout(Ny) && tiled_layout(Nx) ? 1 : 0; }

int namsa_tile_plane(const void* src, void* dst, int Ny, int Nx, void* stream) {
  if (!src || !dst || src == dst) return fail(kErrArg, "namsa_tile_plane: null or aliased planes");
  if (Ny <= 0 || Nx <= 0 || Ny % 2 || Nx % 4)
    return fail(kErrArg, "namsa_tile_plane: %d x %d is not whole 2 x 4 tiles", Ny, Nx);
  const long long n = static_cast<long long>(Ny) * Nx;
  hipLaunchKernelGGL(tile_plane_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), static_cast<const float2*>(src), static_cast<float2*>(dst), Nx, n);
  return check_hip(hipGetLastError(), "tile_plane_kernel launch");
}

int namsa_slice_rows_tiled(void* psi_ws, int B, int Ny, int Nx, const void* t_slice, int band_lo, int band_hi,
                           void* stream) {
  return slice_rows(psi_ws, B, Ny, Nx, t_slice, band_lo, band_hi, 1, stream, "namsa_slice_rows_tiled");
}

int namsa_slice_cols_tiled(void* psi_ws, int B, int Ny, int Nx, const void* pm, int band_lo, int band_hi,
                           void* stream) {
  return slice_cols(psi_ws, B, Ny, Nx, pm, band_lo, band_hi, 1, stream, "namsa_slice_cols_tiled");
}

}  // extern "C"

namespace {

int multislice_impl(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax, const void* t,
                    int S, const void* pm, int band_lo, int band_hi, int group, int mode, float out_scale,
                    void* psi_ws, float* cbed, const float* det, int nDet, double* det_ws, double* det_out,
                    double* mom_ws, double* mom_out, void* stream) {
  const bool pm_tiled = (mode & NAMSA_PM_TILED) != 0;
  mode &= ~NAMSA_PM_TILED;
  if (int rc = check_grid(Ny, Nx, "namsa_multislice")) return rc;
  if (B <= 0 || S < 0 || group < 0 || !psi_k || !pos_yx || !psi_ws)
    return fail(kErrArg, "namsa_multislice: B=%d S=%d group=%d or null probe/positions/workspace",
                B, S, group);
  if (mode != NAMSA_OUT_CBED && mode != NAMSA_OUT_INCIDENT && mode != NAMSA_OUT_EXIT)
    return fail(kErrArg, "namsa_multislice: unknown mode %d", mode);
  if (mode != NAMSA_OUT_INCIDENT && S > 0 && (!t || !pm))
    return fail(kErrArg, "namsa_multislice: null slices or propagator");
  if (mode == NAMSA_OUT_CBED && !cbed) return fail(kErrArg, "namsa_multislice: null cbed");
  if (mode == NAMSA_OUT_CBED && nDet > 0 && (!det || !det_ws || !det_out))
    return fail(kErrArg, "namsa_multislice: null detector buffers");
  if (mom_ws && (mode != NAMSA_OUT_CBED || !mom_out))
    return fail(kErrArg, "namsa_multislice_moments: moments need CBED output and a moments buffer");
  if (int rc = check_band(Nx, band_lo, band_hi, "namsa_multislice")) return rc;

  hipStream_t st = as_stream(stream);
  const float2* tt = static_cast<const float2*>(t);
  const long long plane = static_cast<long long>(Ny) * Nx;
  const float inv_n2 = 1.0f / static_cast<float>(plane);
  const int G = (group > 0 && group < B) ? group : B;
  // Where it pays, the slice loop keeps the wave in the tiled layout
  // (passes.hpp): in CBED mode it never leaves the library, and in exit-wave
  // mode the last (inverse row) pass writes it back row-major in place.
  const bool tiled = (mode == NAMSA_OUT_CBED || mode == NAMSA_OUT_EXIT) && tiled_layout(Nx) && tiled_layout(Ny);
  if (pm_tiled && !tiled)
    return fail(kErrArg, "namsa_multislice: NAMSA_PM_TILED needs CBED or exit-wave mode at a tiled size (not "
                "%d x %d)", Ny, Nx);

  for (int g0 = 0; g0 < B; g0 += G) {
    const int Bg = std::min(G, B - g0);
    // CBED mode: every group reuses the first G planes of the workspace;
    // otherwise the workspace is the output and each group owns its planes.
    float2* psi = static_cast<float2*>(psi_ws) + (mode == NAMSA_OUT_CBED ? 0 : g0 * plane);

    ColArgs a = col_args(psi, Bg, Nx);
    a.psi_k = static_cast<const float2*>(psi_k);
    a.pos = pos_yx + 2LL * g0;
    a.kmax = double(kmax);
    a.pm = static_cast<const float2*>(pm);
    a.pm_scale = inv_n2;
    a.band_lo = band_lo;
    a.band_hi = band_hi;
    a.tiled = tiled ? 1 : 0;
    a.pm_tiled = pm_tiled ? 1 : 0;
    if (int rc = col_pass(Ny, COL_PROBE, a, st)) return rc;

    if (mode == NAMSA_OUT_INCIDENT) {
      if (int rc = row_pass(Nx, ROW_INV, row_args(psi, Bg, Ny, Nx, nullptr, out_scale, 0, Nx), st))
        return rc;
      continue;
    }
    // The probe former's spectrum fills every column; after the first
    // propagator only the band can be nonzero.
    // Each row step's output is read only by the column pass that follows
    // (or the readout), which touches only the column strips meeting the
    // band; strips are at most kStoreAlign wide and aligned to their width,
    // so the row pass stores just the kStoreAlign-aligned band.
    const int last_prop = (mode == NAMSA_OUT_EXIT) ? S : S - 1;
    const bool precise = S > precise_slices(Ny) || S > precise_slices(Nx);
    const int store_lo = band_lo / kStoreAlign * kStoreAlign;
    const int store_hi = std::min(Nx, (band_hi + kStoreAlign - 1) / kStoreAlign * kStoreAlign);
    for (int i = 0; i < S; ++i) {
      const int lo = i == 0 ? 0 : band_lo, hi = i == 0 ? Nx : band_hi;
      RowArgs r = row_args(psi, Bg, Ny, Nx, tt + i * plane, 1.0f, lo, hi);
      r.store_lo = store_lo;
      r.store_hi = store_hi;
      r.tiled = a.tiled;
      if (int rc = row_pass(Nx, ROW_STEP, r, st, precise)) return rc;
      if (i < last_prop)
        if (int rc = col_pass(Ny, COL_PROP, a, st, precise)) return rc;
    }
    if (mode == NAMSA_OUT_EXIT) {
      const int lo = S == 0 ? 0 : band_lo, hi = S == 0 ? Nx : band_hi;
      RowArgs r = row_args(psi, Bg, Ny, Nx, nullptr, out_scale, lo, hi);
      r.tiled = a.tiled;  // reads tiles, writes rows
      if (int rc = row_pass(Nx, ROW_INV, r, st)) return rc;
      continue;
    }
    // CBED readout.  S > 0: the state is u = R(t x) (no 1/N^2 on the row
    // pass), so |sqrt(N^2) FFT(psi_S/N^2)|^2 = |PM C(u)|^2 / N^2.  S == 0: the
    // state is v0 and FFT(psi0) = Nx C(v0), so it is (Nx/Ny) |C(v0)|^2.
    a.apply_pm = S > 0 ? 1 : 0;
    a.amp = S > 0 ? static_cast<float>(out_scale / static_cast<double>(plane))
                  : static_cast<float>(static_cast<double>(Nx) / Ny * out_scale);
    a.cbed = cbed + g0 * plane;
    a.det = det;
    a.n_det = nDet > 0 ? nDet : 0;
    a.det_ws = det_ws;
    a.mom_ws = mom_ws;
    if (int rc = col_pass(Ny, COL_READOUT, a, st, precise)) return rc;
    if (nDet > 0) {
      const int n = Bg * nDet;
      hipLaunchKernelGGL(det_finalize_kernel, dim3((n + 255) / 256), dim3(256), 0, st, det_ws,
                         a.strips_per_image, nDet, Bg, det_out + static_cast<long long>(g0) * nDet);
      if (int rc = check_hip(hipGetLastError(), "det_finalize_kernel launch")) return rc;
    }
    if (mom_ws) {  // the same strip-order sum, two "detectors": the CBED and its square
      const int n = Bg * 2;
      hipLaunchKernelGGL(det_finalize_kernel, dim3((n + 255) / 256), dim3(256), 0, st, mom_ws,
                         a.strips_per_image, 2, Bg, mom_out + 2LL * g0);
      if (int rc = check_hip(hipGetLastError(), "moments finalize launch")) return rc;
    }
  }
  return kOk;
}

}  // namespace

extern "C" {

int namsa_multislice(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                     const void* t, int S, const void* pm, int band_lo, int band_hi, int group,
                     int mode, float out_scale, void* psi_ws, float* cbed, const float* det,
                     int nDet, double* det_ws, double* det_out, void* stream) {
  return multislice_impl(psi_k, pos_yx, B, Ny, Nx, kmax, t, S, pm, band_lo, band_hi, group, mode, out_scale,
                         psi_ws, cbed, det, nDet, det_ws, det_out, nullptr, nullptr, stream);
}

int namsa_multislice_moments(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                             const void* t, int S, const void* pm, int band_lo, int band_hi, int group,
                             float out_scale, void* psi_ws, float* cbed, const float* det, int nDet,
                             double* det_ws, double* det_out, double* mom_ws, double* mom_out, void* stream) {
  if (!mom_ws || !mom_out) return fail(kErrArg, "namsa_multislice_moments: null moments buffers");
  return multislice_impl(psi_k, pos_yx, B, Ny, Nx, kmax, t, S, pm, band_lo, band_hi, group, NAMSA_OUT_CBED,
                         out_scale, psi_ws, cbed, det, nDet, det_ws, det_out, mom_ws, mom_out, stream);
}

int namsa_multislice_noalloc(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                             const void* t, int S, const void* pm, int band_lo, int band_hi, int group,
                             float out_scale, float* cbed, const float* det, int nDet, double* det_out,
                             void* stream) {
  if (int rc = check_grid(Ny, Nx, "namsa_multislice_noalloc")) return rc;
  if (B <= 0 || group < 0 || nDet < 0)
    return fail(kErrArg, "namsa_multislice_noalloc: B=%d group=%d nDet=%d", B, group, nDet);
  // scratch = [G planes of wave state][G x (Nx / narrowest tile) x nDet partial sums]
  const size_t plane_bytes = static_cast<size_t>(Ny) * Nx * sizeof(float2);
  const size_t det_bytes = static_cast<size_t>(Nx / kMinColTile) * nDet * sizeof(double);
  const long long cap = static_cast<long long>(NAMSA_SCRATCH_BYTES / (plane_bytes + det_bytes));
  if (cap < 1)
    return fail(kErrArg, "namsa_multislice_noalloc: one %dx%d probe needs %zu bytes, the scratch holds %zu", Ny, Nx,
                plane_bytes + det_bytes, static_cast<size_t>(NAMSA_SCRATCH_BYTES));
  const long long want = (group > 0 && group < B) ? group : B;
  if (group > 0 && want > cap)
    return fail(kErrArg, "namsa_multislice_noalloc: group %lld exceeds the %lld probes of %dx%d (%d detectors) "
                "the %zu-byte scratch holds", want, cap, Ny, Nx, nDet, static_cast<size_t>(NAMSA_SCRATCH_BYTES));
  const int G = static_cast<int>(std::min(want, cap));
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  void* psi_ws = base;
  double* det_ws = nDet > 0 ? reinterpret_cast<double*>(base + G * plane_bytes) : nullptr;
  return multislice_impl(psi_k, pos_yx, B, Ny, Nx, kmax, t, S, pm, band_lo, band_hi, G, NAMSA_OUT_CBED, out_scale,
                         psi_ws, cbed, det, nDet, det_ws, det_out, nullptr, nullptr, stream);
}

}  // extern "C"
