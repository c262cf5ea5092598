// Pointwise kernels of the coarse-slice corrector network (its matrix
// products run through namsa_gemm_bf16): fills, GELU backward, Adam, the
// spectral modes' bookkeeping, and the spectral corrector's input packing,
// CBED-intensity loss and correction.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "namsa_hip.h"
#include "passes.hpp"  // wave_sum

namespace namsa {

__global__ void fill_kernel(float* __restrict__ x, long long n, float v) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) x[i] = v;
}

__device__ __forceinline__ float gelu_grad_tanh(float x) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (x + 0.044715f * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x * x);
}

__global__ void gelu_backward_kernel(float* __restrict__ g, const float* __restrict__ pre, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x)
    g[i] *= gelu_grad_tanh(pre[i]);
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
                            float bc1, float bc2, float wd) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * ((mi / bc1) / (sqrtf(vi / bc2) + eps) + wd * p[i]);
  }
}

// P[cb][(r,p)][(t,q)] -> Xs[pq][(r,c)][b]; one thread per (pq, c, b)
__global__ void modes_combine_kernel(const float* __restrict__ P, float* __restrict__ Xs, int C, int B,
                                     int m) {
  const long long total = static_cast<long long>(m) * m * C * B;
  const int m2 = 2 * m;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += 256LL * gridDim.x) {
    const int b = static_cast<int>(i % B);
    const int c = static_cast<int>((i / B) % C);
    const int pq = static_cast<int>(i / (static_cast<long long>(B) * C));
    const int p = pq / m, q = pq % m;
    const float* Pc = P + (static_cast<long long>(c) * B + b) * m2 * m2;
    const float p00 = Pc[p * m2 + q], p01 = Pc[p * m2 + m + q];
    const float p10 = Pc[(m + p) * m2 + q], p11 = Pc[(m + p) * m2 + m + q];
    float* X = Xs + static_cast<long long>(pq) * 2 * C * B;
    X[static_cast<long long>(c) * B + b] = p00 + p11;
    X[static_cast<long long>(C + c) * B + b] = p01 - p10;
  }
}

// Ys[pq][(r,o)][b] -> E[ob][(r,p)][(t,q)]; one thread per (o, b, p, q)
__global__ void modes_expand_kernel(const float* __restrict__ Ys, float* __restrict__ E, int C, int B,
                                    int m) {
  const long long total = static_cast<long long>(m) * m * C * B;
  const int m2 = 2 * m;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += 256LL * gridDim.x) {
    const int q = static_cast<int>(i % m);
    const int p = static_cast<int>((i / m) % m);
    const long long ob = i / (static_cast<long long>(m) * m);  // o * B + b
    const int o = static_cast<int>(ob / B), b = static_cast<int>(ob % B);
    const float* Y = Ys + static_cast<long long>(p * m + q) * 2 * C * B;
    const float yr = Y[static_cast<long long>(o) * B + b];
    const float yi = Y[static_cast<long long>(C + o) * B + b];
    float* e = E + ob * m2 * m2;
    e[p * m2 + q] = yr;
    e[(m + p) * m2 + q] = -yi;
    e[p * m2 + m + q] = yi;
    e[(m + p) * m2 + m + q] = yr;
  }
}

// ------------------------------------------------------------------ spectral corrector
// The spectral corrector works on the centred exit spectrum Psi(k) of the
// coarse path (|Psi|^2 is the CBED), inside the window [r0, r0+H) x [c0, c0+W)
// that bounds the bandwidth mask.  Planes are full [B][Ny][Nx]; the network's
// arrays are channel-major [ch][B*H*W] over the window.
struct Window {
  int Ny, Nx, r0, c0, H, W;
};

__device__ __forceinline__ long long plane_index(const Window& w, long long b, int y, int x) {
  return (b * w.Ny + (w.r0 + y)) * static_cast<long long>(w.Nx) + (w.c0 + x);
}

// u[0..1] = s Psi_c, u[2..3] = s Psi_0, u[4] = phase_scale * k^2 (k on sample b's linspace grid of
// extent kmax[b]), u[5] = thick[b] * thick_scale,
// then (opts & 1) log1p(vscale |vhat[vidx[b]]|) (NaN for an index outside [0, nv)), (opts & 2) |s Psi_c|
__global__ void spectral_pack_kernel(const float2* __restrict__ pc, const float2* __restrict__ p0,
                                     const float2* __restrict__ vhat, const int* __restrict__ vidx, int nv, Window w,
                                     int B, const float* __restrict__ kmax, float s, float phase_scale,
                                     const float* __restrict__ thick, float thick_scale, float vscale, int opts,
                                     float* __restrict__ u) {
  const long long hw = static_cast<long long>(w.H) * w.W;
  const long long n = hw * B;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
    const long long b = i / hw;
    const int y = static_cast<int>((i % hw) / w.W), x = static_cast<int>(i % w.W);
    const long long j = plane_index(w, b, y, x);
    const float2 c = pc[j], z = p0[j];
    u[i] = c.x * s;
    u[n + i] = c.y * s;
    u[2 * n + i] = z.x * s;
    u[3 * n + i] = z.y * s;
    const double km = kmax[b];
    const double ky = k_lin(w.r0 + y, km, w.Ny), kx = k_lin(w.c0 + x, km, w.Nx);
    u[4 * n + i] = static_cast<float>(phase_scale * (ky * ky + kx * kx));
    u[5 * n + i] = thick[b] * thick_scale;
    int ch = 6;
    if (opts & 1) {
      const long long pv = vidx ? vidx[b] : 0;
      if (pv >= 0 && pv < nv) {
        const float2 v = vhat[(pv * w.Ny + (w.r0 + y)) * static_cast<long long>(w.Nx) + (w.c0 + x)];
        u[ch * n + i] = log1pf(vscale * sqrtf(v.x * v.x + v.y * v.y));
      } else {
        u[ch * n + i] = __builtin_nanf("");
      }
      ++ch;
    }
    if (opts & 2) u[ch * n + i] = sqrtf(c.x * c.x + c.y * c.y) * s;
  }
}

__global__ void zero_doubles_kernel(double* __restrict__ x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// I = |s Psi_c (1 + o)|^2, F = s^2 target over the window; per sample
// sums[2b] += sum (I - F)^2 and sums[2b+1] += sum F^2 (sums zeroed first).
__global__ void __launch_bounds__(256)
spectral_sums_kernel(const float* __restrict__ out, const float2* __restrict__ pc, const float* __restrict__ tgt,
                     Window w, float s, double* __restrict__ sums) {
  __shared__ double red[2][4];
  const long long hw = static_cast<long long>(w.H) * w.W;
  const long long n = hw * gridDim.y;
  const long long b = blockIdx.y;
  const float s2 = s * s;
  double num = 0.0, den = 0.0;
  for (long long p = blockIdx.x * 256LL + threadIdx.x; p < hw; p += 256LL * gridDim.x) {
    const long long i = b * hw + p;
    const int y = static_cast<int>(p / w.W), x = static_cast<int>(p % w.W);
    const long long j = plane_index(w, b, y, x);
    const float2 c = pc[j];
    const float a = 1.0f + out[i], q = out[n + i];
    const float re = c.x * a - c.y * q, im = c.x * q + c.y * a;
    const float I = (re * re + im * im) * s2, F = tgt[j] * s2;
    num += double(I - F) * double(I - F);
    den += double(F) * double(F);
  }
  num = wave_sum(num);
  den = wave_sum(den);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = num;
    red[1][threadIdx.x >> 6] = den;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&sums[2 * b], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
    atomicAdd(&sums[2 * b + 1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
  }
}

// gout = d/d(o_re, o_im) of mean_b sum (I - F)^2 / sum F^2:
// dI/do_re = 2 s^2 Re(conj(Psi) Psi_c), dI/do_im = -2 s^2 Im(conj(Psi) Psi_c), Psi = Psi_c (1 + o)
__global__ void spectral_grad_kernel(const float* __restrict__ out, const float2* __restrict__ pc,
                                     const float* __restrict__ tgt, Window w, int B, float s,
                                     const double* __restrict__ sums, float* __restrict__ gout) {
  const long long hw = static_cast<long long>(w.H) * w.W;
  const long long n = hw * B;
  const float s2 = s * s;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
    const long long b = i / hw;
    const int y = static_cast<int>((i % hw) / w.W), x = static_cast<int>(i % w.W);
    const long long j = plane_index(w, b, y, x);
    const float2 c = pc[j];
    const float a = 1.0f + out[i], q = out[n + i];
    const float re = c.x * a - c.y * q, im = c.x * q + c.y * a;
    const float I = (re * re + im * im) * s2, F = tgt[j] * s2;
    const float g = static_cast<float>(2.0 / ((sums[2 * b + 1] + 1e-300) * B)) * (I - F) * 2.0f * s2;
    // conj(Psi) Psi_c = (re - i im)(c.x + i c.y)
    const float pr = re * c.x + im * c.y, pi = re * c.y - im * c.x;
    gout[i] = g * pr;
    gout[n + i] = -g * pi;
  }
}

// cbed = |Psi_c (1 + o)|^2 * scale inside the window, |Psi_c|^2 * scale outside
__global__ void spectral_apply_kernel(const float* __restrict__ out, const float2* __restrict__ pc, Window w, int B,
                                      float scale, float* __restrict__ cbed) {
  const long long plane = static_cast<long long>(w.Ny) * w.Nx;
  const long long hw = static_cast<long long>(w.H) * w.W;
  const long long n = plane * B, nw = hw * B;
  for (long long j = blockIdx.x * 256LL + threadIdx.x; j < n; j += 256LL * gridDim.x) {
    const long long b = j / plane;
    const int y = static_cast<int>((j % plane) / w.Nx) - w.r0, x = static_cast<int>(j % w.Nx) - w.c0;
    const float2 c = pc[j];
    float re = c.x, im = c.y;
    if (y >= 0 && y < w.H && x >= 0 && x < w.W) {
      const long long i = (b * w.H + y) * w.W + x;
      const float a = 1.0f + out[i], q = out[nw + i];
      re = c.x * a - c.y * q;
      im = c.x * q + c.y * a;
    }
    cbed[j] = (re * re + im * im) * scale;
  }
}

// out = (scale v (-1)^(y+x) if checker else scale v, 0): a real plane as a
// complex one whose forward transform comes out centred
__global__ void real_to_complex_kernel(const float* __restrict__ v, long long n, int Ny, int Nx, float scale,
                                       int checker, float2* __restrict__ out) {
  const long long plane = static_cast<long long>(Ny) * Nx;
  for (long long j = blockIdx.x * 256LL + threadIdx.x; j < n; j += 256LL * gridDim.x) {
    float a = v[j] * scale;
    if (checker && ((((j % plane) / Nx) + (j % Nx)) & 1)) a = -a;
    out[j] = make_float2(a, 0.0f);
  }
}

inline int ew_grid(long long n) {
  return static_cast<int>(std::min<long long>((n + 255) / 256, 65536));
}

int check_window(int B, int Ny, int Nx, int r0, int c0, int H, int W, const char* who) {
  if (B <= 0 || B > 65535 || Ny <= 0 || Nx <= 0 || H <= 0 || W <= 0 || r0 < 0 || c0 < 0 || r0 + H > Ny ||
      c0 + W > Nx)
    return fail(kErrArg, "%s: window [%d,%d)x[%d,%d) outside %dx%d or B=%d", who, r0, r0 + H, c0, c0 + W, Ny, Nx,
                B);
  return kOk;
}

}  // namespace namsa

using namespace namsa;

extern "C" {

int namsa_fill(float* x, long long n, float v, void* stream) {
  if (!x || n < 0) return fail(kErrArg, "namsa_fill: bad arguments");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(fill_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream), x, n, v);
  return check_hip(hipGetLastError(), "fill_kernel launch");
}

int namsa_gelu_backward(float* g, const float* pre, long long n, void* stream) {
  if (!g || !pre || n <= 0) return fail(kErrArg, "namsa_gelu_backward: bad arguments");
  hipLaunchKernelGGL(gelu_backward_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream), g, pre, n);
  return check_hip(hipGetLastError(), "gelu_backward_kernel launch");
}

int namsa_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
               float eps, int step, float wd, void* stream) {
  if (!p || !g || !m || !v || n <= 0 || step < 1) return fail(kErrArg, "namsa_adam: bad arguments");
  const float bc1 = 1.0f - powf(b1, float(step)), bc2 = 1.0f - powf(b2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream), p, g, m, v, n, lr,
                     b1, b2, eps, bc1, bc2, wd);
  return check_hip(hipGetLastError(), "adam_kernel launch");
}

int namsa_modes_combine(const float* P, float* Xs, int C, int B, int m, void* stream) {
  if (!P || !Xs || C <= 0 || B <= 0 || m <= 0) return fail(kErrArg, "namsa_modes_combine: bad arguments");
  const long long n = static_cast<long long>(m) * m * C * B;
  hipLaunchKernelGGL(modes_combine_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream), P, Xs, C, B, m);
  return check_hip(hipGetLastError(), "modes_combine_kernel launch");
}

int namsa_modes_expand(const float* Ys, float* E, int C, int B, int m, void* stream) {
  if (!Ys || !E || C <= 0 || B <= 0 || m <= 0) return fail(kErrArg, "namsa_modes_expand: bad arguments");
  const long long n = static_cast<long long>(m) * m * C * B;
  hipLaunchKernelGGL(modes_expand_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream), Ys, E, C, B, m);
  return check_hip(hipGetLastError(), "modes_expand_kernel launch");
}

int namsa_spectral_pack(const void* psi_c, const void* psi_0, const void* vhat, const int* vidx, int nv, int B, int Ny,
                        int Nx, int r0, int c0, int H, int W, const float* kmax, float s, float phase_scale,
                        const float* thick, float thick_scale, float vscale, int opts, float* u, void* stream) {
  if (int rc = check_window(B, Ny, Nx, r0, c0, H, W, "namsa_spectral_pack")) return rc;
  if (!psi_c || !psi_0 || !kmax || !thick || !u || ((opts & 1) && (!vhat || nv < 1)) || opts < 0 || opts > 3)
    return fail(kErrArg, "namsa_spectral_pack: null input or opts=%d", opts);
  const long long n = static_cast<long long>(B) * H * W;
  hipLaunchKernelGGL(spectral_pack_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream),
                     static_cast<const float2*>(psi_c), static_cast<const float2*>(psi_0),
                     static_cast<const float2*>(vhat), vidx, nv, Window{Ny, Nx, r0, c0, H, W}, B, kmax, s,
                     phase_scale, thick, thick_scale, vscale, opts, u);
  return check_hip(hipGetLastError(), "spectral_pack_kernel launch");
}

int namsa_spectral_loss(const float* out, const void* psi_c, const float* target, int B, int Ny, int Nx, int r0,
                        int c0, int H, int W, float s, double* sums, float* gout, void* stream) {
  if (int rc = check_window(B, Ny, Nx, r0, c0, H, W, "namsa_spectral_loss")) return rc;
  if (!out || !psi_c || !target || !sums) return fail(kErrArg, "namsa_spectral_loss: null argument");
  hipStream_t st = as_stream(stream);
  const Window w{Ny, Nx, r0, c0, H, W};
  const float2* pc = static_cast<const float2*>(psi_c);
  hipLaunchKernelGGL(zero_doubles_kernel, dim3((2 * B + 255) / 256), dim3(256), 0, st, sums, 2 * B);
  if (int rc = check_hip(hipGetLastError(), "zero_doubles_kernel launch")) return rc;
  const long long hw = static_cast<long long>(H) * W;
  const unsigned chunks = static_cast<unsigned>(std::min<long long>((hw + 255) / 256, 1024));
  hipLaunchKernelGGL(spectral_sums_kernel, dim3(chunks, B), dim3(256), 0, st, out, pc, target, w, s, sums);
  if (int rc = check_hip(hipGetLastError(), "spectral_sums_kernel launch")) return rc;
  if (!gout) return kOk;
  hipLaunchKernelGGL(spectral_grad_kernel, dim3(ew_grid(hw * B)), dim3(256), 0, st, out, pc, target, w, B, s, sums,
                     gout);
  return check_hip(hipGetLastError(), "spectral_grad_kernel launch");
}

int namsa_spectral_apply(const float* out, const void* psi_c, int B, int Ny, int Nx, int r0, int c0, int H, int W,
                         float scale, float* cbed, void* stream) {
  if (int rc = check_window(B, Ny, Nx, r0, c0, H, W, "namsa_spectral_apply")) return rc;
  if (!out || !psi_c || !cbed) return fail(kErrArg, "namsa_spectral_apply: null argument");
  const long long n = static_cast<long long>(B) * Ny * Nx;
  hipLaunchKernelGGL(spectral_apply_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream), out,
                     static_cast<const float2*>(psi_c), Window{Ny, Nx, r0, c0, H, W}, B, scale, cbed);
  return check_hip(hipGetLastError(), "spectral_apply_kernel launch");
}

int namsa_real_to_complex(const float* v, int B, int Ny, int Nx, float scale, int checker, void* out,
                          void* stream) {
  if (!v || !out || B <= 0 || Ny <= 0 || Nx <= 0) return fail(kErrArg, "namsa_real_to_complex: bad arguments");
  const long long n = static_cast<long long>(B) * Ny * Nx;
  hipLaunchKernelGGL(real_to_complex_kernel, dim3(ew_grid(n)), dim3(256), 0, as_stream(stream), v, n, Ny, Nx, scale,
                     checker ? 1 : 0, static_cast<float2*>(out));
  return check_hip(hipGetLastError(), "real_to_complex_kernel launch");
}

}  // extern "C"
