// Channel-last BF16 inference of the corrector network on the matrix cores.
//
// The network's activations are 32 channels per pixel; here they live as
// bf16 [rows][W][32] (rows = B * H, one image row of one sample each), so one
// image row is a W x 32 matrix and every product of a Fourier layer that
// touches the full activation is a row-local MFMA product:
//
//   x_{l+1}[x][o] = gelu( alpha sum_q Fx[x][q] Zp_l[o][q] + sum_c x_l[x][c] W_l[o][c] + b_l[o] )
//   Z_{l+1}[c][q] = sum_x x_{l+1}[x][c] Fx[x][q]
//
// with Fx[x][q] the truncated row DFT ([cos | -sin] of the m retained x modes,
// 2m = 32 columns), Zp_l the row-inverse operand the small mode-space
// products leave (mode mixing and the y transforms stay on the strided GEMM),
// and Z_{l+1} the next layer's row DFT.  One wavefront streams one row in
// 32-pixel tiles: the first product is four v_mfma_f32_32x32x16_bf16 with
// K = 32 modes + 32 channels (both operands read 16 bytes per lane straight
// from memory), the tile goes through LDS once (bias, GELU, bf16) and leaves
// as the next activation, and two more MFMAs fold it into the next row DFT.
// Each layer therefore reads and writes the activation once: 128 bytes per
// pixel.  The lift (features -> 32 channels) also forms layer 0's row DFT;
// the last layer feeds the pointwise head (32 -> 64 -> 2) instead.
//
// MFMA 32x32x16 operand layout (lane l, 8 bf16 per operand register set):
//   A: row l%32, k = 8*(l/32) + e      B: column l%32, k = 8*(l/32) + e
//   D: 16 floats, register r holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
// Z / Zp rows are ordered (c, b, y): row (c * B + b) * H + y, 32 floats each,
// the layout of the strided GEMM path's row-DFT output.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "namsa_hip.h"

namespace namsa {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kCh = 32;        // channels
constexpr int kQ = 32;         // 2m row-DFT columns
constexpr int kHid = 64;       // head width
constexpr int kTile = 32;      // pixels per tile
constexpr int kPitch = 40;     // LDS row pitch in bf16 (80 bytes: 16-byte aligned rows)
constexpr int kWaves = 4;

// tanh-GELU, x/2 (1 + tanh(u)) = x (1 - 1 / (exp(2u) + 1)), u = sqrt(2/pi) (x + 0.044715 x^3);
// the exponential form differs from tanhf by ~1e-7, far below the bf16 rounding of its result
__device__ __forceinline__ float gelu_t(float x) {
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  return x - x / (__expf(2.0f * u) + 1.0f);
}

// The LDS tiles are private to a wavefront: its lanes exchange through them
// in program order, so a wavefront-scope fence (no wait on other wavefronts)
// is the only ordering the exchanges need.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ bf16x8 cvt8(const float* p, float scale) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  const float4 b = *reinterpret_cast<const float4*>(p + 4);
  bf16x8 v;
  v[0] = __bf16(a.x * scale);
  v[1] = __bf16(a.y * scale);
  v[2] = __bf16(a.z * scale);
  v[3] = __bf16(a.w * scale);
  v[4] = __bf16(b.x * scale);
  v[5] = __bf16(b.y * scale);
  v[6] = __bf16(b.z * scale);
  v[7] = __bf16(b.w * scale);
  return v;
}

struct RowGeom {
  int rows, W, Wp, H, B;
};

// The row DFT of one 32-pixel tile held transposed in LDS (tt[c][x]) added to
// zacc(q, c): A(q, k = x) = fxT[q][x0 + k], B(k = x, c) = tt[c][k].
__device__ __forceinline__ void tile_row_dft(f32x16& zacc, const __bf16* __restrict__ fxT, int Wp, int x0,
                                             const __bf16 (*tt)[kPitch], int lane) {
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int k = 16 * s + 8 * (lane >> 5);
    const bf16x8 a = ld8(fxT + static_cast<long long>(lane & 31) * Wp + x0 + k);
    const bf16x8 b = ld8(&tt[lane & 31][k]);
    zacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, zacc, 0, 0, 0);
  }
}

__device__ __forceinline__ void store_z(float* __restrict__ Z, const f32x16& zacc, const RowGeom& g, int r,
                                        int lane) {
  const int b = r / g.H, y = r % g.H;
  const int c = lane & 31;
  float* zr = Z + ((static_cast<long long>(c) * g.B + b) * g.H + y) * kQ;
#pragma unroll
  for (int i = 0; i < 16; ++i) zr[8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)] = zacc[i];
}

// x0 = Wl u + bl (u channel-major [cin][rows * W]) and layer 0's row DFT.
// Per 32-pixel tile one MFMA with K = 16 (cin inputs, the rest zero):
// A(x, k) = u[k][x], B(k, o) = Wl[o][k].
__global__ void __launch_bounds__(256)
fno_lift_kernel(const float* __restrict__ u, int cin, const float* __restrict__ Wl, const float* __restrict__ bl,
                const __bf16* __restrict__ fxT, RowGeom g, __bf16* __restrict__ xout, float* __restrict__ Z) {
  __shared__ __align__(16) __bf16 tt_all[kWaves][kCh][kPitch];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = blockIdx.x * kWaves + wave;
  const bool live = r < g.rows;
  __bf16 (*tt)[kPitch] = tt_all[wave];
  const long long n = static_cast<long long>(g.rows) * g.W;
  const int col = lane & 31, kh = 8 * (lane >> 5);
  bf16x8 wop;                                    // B(k = kh + e, o = col)
#pragma unroll
  for (int e = 0; e < 8; ++e) wop[e] = __bf16(kh + e < cin ? Wl[col * cin + kh + e] : 0.0f);
  const float bo = bl[col];
  f32x16 zacc;
#pragma unroll
  for (int i = 0; i < 16; ++i) zacc[i] = 0.0f;
  for (int x0 = 0; x0 < g.W; x0 += kTile) {
    const int pa = x0 + col;                     // this lane's A row (pixel)
    const bool va = live && pa < g.W;
    bf16x8 a;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      a[e] = __bf16((va && kh + e < cin) ? u[(kh + e) * n + static_cast<long long>(r) * g.W + pa] : 0.0f);
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, wop, acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int x = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
      tt[col][x] = __bf16(x0 + x < g.W ? acc[i] + bo : 0.0f);
    }
    wave_sync();
    const int px = x0 + col;                     // store: lane -> pixel l%32, channels 16(l/32) .. +16
    if (live && px < g.W) {
      bf16x8 o0, o1;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        o0[e] = tt[16 * (lane >> 5) + e][col];
        o1[e] = tt[16 * (lane >> 5) + 8 + e][col];
      }
      __bf16* dst = xout + (static_cast<long long>(r) * g.W + px) * kCh + 16 * (lane >> 5);
      *reinterpret_cast<bf16x8*>(dst) = o0;
      *reinterpret_cast<bf16x8*>(dst + 8) = o1;
    }
    tile_row_dft(zacc, fxT, g.Wp, x0, tt, lane);
    wave_sync();
  }
  if (live) store_z(Z, zacc, g, r, lane);
}

// One Fourier layer.  last == 0: x_out = gelu(...) and Z_next; last == 1: no
// GELU, then the head out[t][p] = Q2 gelu(Q1 x + bq1) + bq2 (out channel-major
// [2][rows * W] float).
template <int LAST>
__global__ void __launch_bounds__(256)
fno_layer_kernel(const __bf16* __restrict__ x_in, const float* __restrict__ Zp, const __bf16* __restrict__ fx,
                 const __bf16* __restrict__ fxT, const float* __restrict__ Wt, const float* __restrict__ bias,
                 float alpha, RowGeom g, __bf16* __restrict__ x_out, float* __restrict__ Z_next,
                 const float* __restrict__ Q1, const float* __restrict__ bq1, const float* __restrict__ Q2,
                 const float* __restrict__ bq2, float* __restrict__ out) {
  __shared__ __align__(16) __bf16 tt_all[kWaves][kCh][kPitch];           // [c][x] (LAST: [x][c])
  constexpr int kHPitch = kHid + 8;                                      // 144-byte rows: 16-byte aligned
  __shared__ __align__(16) __bf16 hb_all[LAST ? kWaves : 1][LAST ? kTile : 1][LAST ? kHPitch : 8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = blockIdx.x * kWaves + wave;
  const bool live = r < g.rows;
  const int rr = live ? r : 0;                  // masked waves read row 0 and write nothing
  __bf16 (*tt)[kPitch] = tt_all[wave];
  const int b = rr / g.H, y = rr % g.H;
  const int col = lane & 31, kh = 8 * (lane >> 5);

  // B operands of the layer product, the same for every tile of the row:
  // k < 32: alpha Zp[o][q] (q = k), k >= 32: W[o][c] (c = k - 32), o = column
  bf16x8 bop[4];
  const float* zp_row = Zp + ((static_cast<long long>(col) * g.B + b) * g.H + y) * kQ;
  bop[0] = cvt8(zp_row + kh, alpha);
  bop[1] = cvt8(zp_row + 16 + kh, alpha);
  bop[2] = cvt8(Wt + col * kCh + kh, 1.0f);
  bop[3] = cvt8(Wt + col * kCh + 16 + kh, 1.0f);
  const float bo = bias[col];
  bf16x8 q1op[2][2];                            // head operands B(k = c, j) = Q1[j][c], constant per wave
  bf16x8 q2op[4];                               // B(k = j, t) = Q2[t][j] for t = col < 2, else 0
  if (LAST) {
#pragma unroll
    for (int jb = 0; jb < 2; ++jb)
#pragma unroll
      for (int s = 0; s < 2; ++s) q1op[jb][s] = cvt8(Q1 + (32 * jb + col) * kCh + 16 * s + kh, 1.0f);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int e = 0; e < 8; ++e) q2op[s][e] = __bf16(col < 2 ? Q2[col * kHid + 16 * s + kh + e] : 0.0f);
  }

  f32x16 zacc;
#pragma unroll
  for (int i = 0; i < 16; ++i) zacc[i] = 0.0f;
  const long long n = static_cast<long long>(g.rows) * g.W;
  // channel operands A(x, k = 32 + 16 s + kh ..) of a tile; the next tile's are loaded while this one computes
  auto load_ch = [&](int x0, bf16x8 (&a)[2]) {
    const int pa = x0 + col;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (pa < g.W) {
        a[s] = ld8(x_in + (static_cast<long long>(rr) * g.W + pa) * kCh + 16 * s + kh);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) a[s][e] = __bf16(0.0f);
      }
    }
  };
  bf16x8 anext[2];
  load_ch(0, anext);
  for (int x0 = 0; x0 < g.W; x0 += kTile) {
    const int pa = x0 + col;                    // this lane's A row (pixel)
    bf16x8 ach[2] = {anext[0], anext[1]};
    if (x0 + kTile < g.W) load_ch(x0 + kTile, anext);
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
#pragma unroll
    for (int s = 0; s < 2; ++s) {               // k = 16 s + kh .. : modes
      const bf16x8 a = ld8(fx + static_cast<long long>(pa) * kQ + 16 * s + kh);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bop[s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)                 // k = 32 + 16 s + kh .. : channels
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ach[s], bop[2 + s], acc, 0, 0, 0);
    // epilogue: pixel x = 8(i/4) + 4(l/32) + i%4 of the tile, channel o = col
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int x = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
      float v = acc[i] + bo;
      if (!LAST) v = gelu_t(v);
      if (x0 + x >= g.W) v = 0.0f;
      if (LAST) tt[x][col] = __bf16(v);
      else tt[col][x] = __bf16(v);
    }
    wave_sync();
    if constexpr (!LAST) {
      // the tile leaves as the next activation: lane -> pixel l%32, channels 16(l/32) .. +16
      const int px = x0 + col;
      if (live && px < g.W) {
        bf16x8 o0, o1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          o0[j] = tt[16 * (lane >> 5) + j][col];
          o1[j] = tt[16 * (lane >> 5) + 8 + j][col];
        }
        __bf16* dst = x_out + (static_cast<long long>(r) * g.W + px) * kCh + 16 * (lane >> 5);
        *reinterpret_cast<bf16x8*>(dst) = o0;
        *reinterpret_cast<bf16x8*>(dst + 8) = o1;
      }
      tile_row_dft(zacc, fxT, g.Wp, x0, tt, lane);
    } else {
      // head: h = gelu(Q1 x + bq1) as two 32-wide MFMA blocks, A(x, k = c) = tt[x][c], kept bf16 in LDS
      __bf16 (*hb)[kHPitch] = hb_all[wave];
#pragma unroll
      for (int jb = 0; jb < 2; ++jb) {
        f32x16 hacc;
#pragma unroll
        for (int i = 0; i < 16; ++i) hacc[i] = 0.0f;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const bf16x8 a = ld8(&tt[col][16 * s + kh]);
          hacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, q1op[jb][s], hacc, 0, 0, 0);
        }
        const float bj = bq1[32 * jb + col];
#pragma unroll
        for (int i = 0; i < 16; ++i)
          hb[8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)][32 * jb + col] = __bf16(gelu_t(hacc[i] + bj));
      }
      wave_sync();
      // out[t][p] = Q2[t] . h + bq2[t]: four MFMAs with K = 64, A(x, k = j) = hb[x][j]; columns t < 2 kept
      f32x16 oacc;
#pragma unroll
      for (int i = 0; i < 16; ++i) oacc[i] = 0.0f;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bf16x8 a = ld8(&hb[col][16 * s + kh]);
        oacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, q2op[s], oacc, 0, 0, 0);
      }
      if (live && col < 2) {
        const float bt = bq2[col];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int px = x0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
          if (px < g.W) out[col * n + static_cast<long long>(r) * g.W + px] = oacc[i] + bt;
        }
      }
    }
    wave_sync();
  }
  if constexpr (!LAST) {
    if (live) store_z(Z_next, zacc, g, r, lane);
  }
}

int check_geom(int B, int H, int W, int Wp, const char* who) {
  if (B <= 0 || H <= 0 || W <= 0 || Wp < W || Wp % kTile != 0 || static_cast<long long>(B) * H > 0x7fffffffLL)
    return fail(kErrArg, "%s: bad geometry B=%d H=%d W=%d Wp=%d (Wp must be a multiple of 32, >= W)", who, B, H,
                W, Wp);
  return kOk;
}

}  // namespace
}  // namespace namsa

using namespace namsa;

extern "C" {

int namsa_fno_lift(const float* u, int cin, const float* Wl, const float* bl, const void* fxT, int B, int H, int W,
                   int Wp, void* x0, float* Z, void* stream) {
  if (int rc = check_geom(B, H, W, Wp, "namsa_fno_lift")) return rc;
  if (!u || !Wl || !bl || !fxT || !x0 || !Z || cin < 1 || cin > 8)
    return fail(kErrArg, "namsa_fno_lift: null argument or cin=%d outside 1..8", cin);
  const RowGeom g{B * H, W, Wp, H, B};
  const unsigned blocks = static_cast<unsigned>((g.rows + kWaves - 1) / kWaves);
  hipLaunchKernelGGL(fno_lift_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), u, cin, Wl, bl,
                     static_cast<const __bf16*>(fxT), g, static_cast<__bf16*>(x0), Z);
  return check_hip(hipGetLastError(), "fno_lift_kernel launch");
}

int namsa_fno_layer(const void* x_in, const float* Zp, const void* fx, const void* fxT, const float* Wt,
                    const float* bias, float alpha, int B, int H, int W, int Wp, void* x_out, float* Z_next,
                    const float* Q1, const float* bq1, const float* Q2, const float* bq2, float* out, void* stream) {
  if (int rc = check_geom(B, H, W, Wp, "namsa_fno_layer")) return rc;
  if (!x_in || !Zp || !fx || !fxT || !Wt || !bias) return fail(kErrArg, "namsa_fno_layer: null operand");
  const bool last = out != nullptr;
  if (last ? (!Q1 || !bq1 || !Q2 || !bq2) : (!x_out || !Z_next))
    return fail(kErrArg, "namsa_fno_layer: a layer needs x_out and Z_next, the last layer Q1, bq1, Q2, bq2 and out");
  const RowGeom g{B * H, W, Wp, H, B};
  const unsigned blocks = static_cast<unsigned>((g.rows + kWaves - 1) / kWaves);
  const auto* xi = static_cast<const __bf16*>(x_in);
  const auto* f = static_cast<const __bf16*>(fx);
  const auto* ft = static_cast<const __bf16*>(fxT);
  if (last)
    hipLaunchKernelGGL(fno_layer_kernel<1>, dim3(blocks), dim3(256), 0, as_stream(stream), xi, Zp, f, ft, Wt, bias,
                       alpha, g, nullptr, nullptr, Q1, bq1, Q2, bq2, out);
  else
    hipLaunchKernelGGL(fno_layer_kernel<0>, dim3(blocks), dim3(256), 0, as_stream(stream), xi, Zp, f, ft, Wt, bias,
                       alpha, g, static_cast<__bf16*>(x_out), Z_next, nullptr, nullptr, nullptr, nullptr, nullptr);
  return check_hip(hipGetLastError(), "fno_layer_kernel launch");
}

}  // extern "C"
