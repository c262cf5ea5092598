// Kernel templates of the row and column passes, parameterised by a schedule.
//
// A row schedule fixes the radix list and elements per thread E of the
// x-transforms; rows of length N are handled by T = N/E threads, several rows
// per 256-thread workgroup when T < 256.  A column schedule also fixes the
// tile width C (adjacent columns per workgroup, C*8 contiguous bytes per row
// segment).  PF prefetches the next row/tile of the persistent loop into
// registers while the current one is transformed, so each wavefront keeps a
// load stream in flight through its compute phase.  NT marks the wave-state
// loads and stores non-temporal so the streamed state does not evict the
// shared slice and propagator from the caches.
#pragma once
#include <algorithm>
#include <type_traits>

#include "fft.hpp"
#include "passes.hpp"

namespace namsa {

constexpr int kRowThreads = 256;

typedef float f2v __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ float2 ld(const float2* p) {
  if constexpr (NT) {
    const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
    return make_float2(v.x, v.y);
  } else {
    return *p;
  }
}

// Wave-state stores.  The default form is a relaxed store at device scope
// (global_store ... sc1): the next pass reads the state from other XCDs, and
// with device-scope stores the L2 write-back at the end of each launch has
// less left to flush -- measured 8.3 -> 6.4 us of fixed cost per slice step
// at 512^2, about 2% of the slice loop.  Relaxed: ordering still comes from
// the launch boundary.
template <bool NT>
__device__ __forceinline__ void st(float2* p, float2 v) {
  if constexpr (NT) {
    f2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
  } else {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __builtin_bit_cast(unsigned long long, v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// W > 0 asks the compiler for at least W wavefronts per SIMD (a register
// budget of 512 / W VGPRs): the row pass keeps most of its memory traffic
// under its arithmetic only with enough wavefronts in flight.
// TR: the slice step holds each transmission row across the images of the
// group (row_step_tr_kernel) instead of reading it once per image.  LS: LDS
// addressed as a lane base plus constant offsets where the padding allows
// (RowLdsT<true>).
template <int N_, int E_, class RL_, bool PF_, bool NT_, bool CT_ = false, int W_ = 0, bool TR_ = false,
          bool LS_ = false>
struct RowS {
  static_assert(!(PF_ && TR_), "the transmission-resident step does not prefetch");
  static constexpr int N = N_, E = E_, W = W_;
  using RL = RL_;
  static constexpr bool PF = PF_, NT = NT_, CT = CT_, TR = TR_, LS = LS_;
};

// Fills the LDS quarter table of N-th roots an LR transform reads its
// twiddles from, and points the LDS accessor at it (workgroup barrier).
// Column schedules only: the row passes are VALU-bound, and forming the
// twiddles there costs more than the registers it frees (5-10% slower at
// 2048^2).
template <int N, class Lds>
__device__ __forceinline__ void load_roots(float2* roots, Lds& lds, int threads) {
  for (int i = threadIdx.x; i < N / 4; i += threads) roots[i] = make_float2(kRootsQ<N>.v[i].c, kRootsQ<N>.v[i].s);
  __syncthreads();
  lds.roots = roots;
}

// Column schedules also take W (a register budget for W wavefronts per SIMD,
// e.g. two 512-thread workgroups per CU at W = 4; the launcher then gives each
// workgroup one (strip, image) task), PG (PROP reads the propagator at each
// use instead of holding the strip's values in registers) and SB (scalar
// bases: element (jt + T m, col) is addressed as a wavefront-uniform base --
// plane plus m T rows, in scalar registers -- plus one 32-bit lane offset,
// instead of E 64-bit vector addresses that the compiler keeps live across
// the persistent loop: 64 VGPRs at E = 16 for the loads and the stores).
template <int N_, int E_, class RL_, int C_, bool PF_, bool NT_, bool DB_ = false, bool CT_ = false, int W_ = 0,
          bool PG_ = false, bool SB_ = false, int K_ = 1, bool LR_ = false>
struct ColS {
  static_assert(C_ >= kMinColTile && C_ <= kMaxColTile, "column tile width out of range");
  static_assert(K_ == 1 || W_ > 0, "images per workgroup apply to the multi-workgroup launch");
  static constexpr int N = N_, E = E_, C = C_, W = W_, K = K_;
  using RL = RL_;
  static constexpr bool PF = PF_, NT = NT_, DB = DB_, CT = CT_, PG = PG_, SB = SB_, LR = LR_;
  static constexpr int threads() { return C * (N / E); }
};

// Raw buffer access to one [Ny][Nx] complex64 plane: the resource (plane
// base) and the byte offset of a row block are wavefront-uniform and live in
// scalar registers, the lane's byte offset in one VGPR.  Written as buffer
// instructions the compiler keeps this form; with plain pointers it folds the
// uniform parts into per-element 64-bit vector addresses.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kBufDword3 = 0x00020000;  // gfx9 raw-buffer resource word 3
constexpr int kBufNt = 2, kBufSc1 = 16;  // cache-policy bits of the aux operand

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float2* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(base), 0, static_cast<int>(bytes), kBufDword3);
}
template <bool NT>
__device__ __forceinline__ float2 buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, NT ? kBufNt : 0));
}
// NT: non-temporal; otherwise sc1, the device-scope store of st<false>
template <bool NT>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, NT ? kBufNt : kBufSc1);
}

// exp(2 pi i * turns) with the phase reduced to [-1/2, 1/2] turns in double.
__device__ __forceinline__ float2 unit_phase(double turns) {
  const double f = turns - rint(turns);
  double s, c;
  sincospi(2.0 * f, &s, &c);
  return make_float2(float(c), float(s));
}

// ------------------------------------------------------------------ row pass
// TL: the wave is stored in 2 x 4 tiles (tile_offset); rows r = 2k, 2k + 1 of
// the stacked images form tile row k, and element t + T m of a row sits at a
// lane base plus 2 T m.  The inverse row transform (ROW_INV, the exit wave)
// reads tiles and writes rows: tile row k and rows 2k, 2k + 1 occupy the same
// 2 Nx elements, which one workgroup holds in registers, so the conversion
// runs in place.  The two rows of a workgroup are interleaved over the
// lanes (row g = lane % 2), so that each wavefront access covers whole
// tiles: with device-scope stores, half-tile pieces written by two
// wavefronts took 40% longer than whole tiles (2048^2, 6 images: a
// memory-only row pass 69.5 against 49.2 us).  Skewing the second row's LDS
// buffer against bank conflicts between the rows made no difference.
template <class S, bool TL>
struct RowMap {
  static constexpr int T = S::N / S::E, G = kRowThreads / T;
  static_assert(!TL || G == 2, "tiled rows pair two rows per workgroup");
  __device__ static int g() { return TL ? threadIdx.x % G : threadIdx.x / T; }
  __device__ static int t() { return TL ? threadIdx.x / G : threadIdx.x % T; }
};

template <class S, int OP, bool TL = false>
__global__ void __launch_bounds__(kRowThreads, S::W > 0 ? S::W : 1) row_kernel(RowArgs a) {
  constexpr int N = S::N, E = S::E;
  using F = Fft<N, E, typename S::RL, S::CT>;
  constexpr int T = F::T;
  constexpr int G = kRowThreads / T;
  constexpr bool STEP = OP == ROW_STEP;
  using Sync = typename std::conditional<(T <= 64 && !TL), WaveSync, BlockSync>::type;
  using Lds = RowLdsT<S::LS>;
  using Map = RowMap<S, TL>;
  __shared__ float2 smem[G * Lds::padded(N)];

  const int g = Map::g();
  const int t = Map::t();
  Lds lds{smem + g * Lds::padded(N)};
  float2 tw[F::NTW_ALLOC];
  F::twiddles(t, tw);

  const long long n_blocks = (a.n_rows + G - 1) / G;
  auto clamp_row = [&](long long rb) {
    const long long r = rb * G + g;
    return r < a.n_rows ? r : a.n_rows - 1;
  };
  // Row of the slice for the next row loaded (row mod rows_per_image), kept
  // incrementally: a 64-bit modulo per row costs ~40 VALU instructions.  Only
  // a clamped tail row (never stored) can get another in-range slice row.
  const int rows_img = a.rows_per_image;
  const int row_step = static_cast<int>((static_cast<long long>(gridDim.x) * G) % rows_img);
  int trow = static_cast<int>(clamp_row(blockIdx.x) % rows_img);
  auto next_trow = [&]() {
    trow += row_step;
    if (trow >= rows_img) trow -= rows_img;
  };
  constexpr int XS = TL ? 2 * T : T;  // element stride of a row in the wave layout
  constexpr bool TL_OUT = TL && OP != ROW_INV;  // stores in the tiled layout too
  constexpr int XS_OUT = TL_OUT ? 2 * T : T;
  static_assert(!TL || T % 4 == 0, "tiled rows need whole tiles per thread group");
  auto row_base = [&](long long r) { return TL ? a.psi + tile_offset(r, t, N) : a.psi + r * N + t; };
  auto out_base = [&](long long r) { return TL_OUT ? a.psi + tile_offset(r, t, N) : a.psi + r * N + t; };
  auto load = [&](long long r, float2(&v)[E], float2(&tv)[E]) {
    const float2* p = row_base(r);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int x = t + T * m;
      v[m] = (x >= a.band_lo && x < a.band_hi) ? ld<S::NT>(p + XS * m) : make_float2(0.f, 0.f);
    }
    if constexpr (STEP) {
      const float2* tr = a.tsl + trow * N + t;
#pragma unroll
      for (int m = 0; m < E; ++m) tv[m] = tr[T * m];
    }
  };

  float2 pv[E], pt[E];
  long long rb = blockIdx.x;
  if constexpr (S::PF) {
    if (rb < n_blocks) load(clamp_row(rb), pv, pt);
  }
  for (; rb < n_blocks; rb += gridDim.x) {
    const long long row = rb * G + g;
    const long long rr = clamp_row(rb);
    float2 v[E], tv[E];
    if constexpr (S::PF) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        v[m] = pv[m];
        tv[m] = pt[m];
      }
      next_trow();
      if (rb + gridDim.x < n_blocks) load(clamp_row(rb + gridDim.x), pv, pt);
    } else {
      load(rr, v, tv);
      next_trow();
    }
    if constexpr (STEP) {
      F::template run<true, Sync>(v, lds, t, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(v[m], tv[m]);
      F::template run<false, Sync>(v, lds, t, tw);
    } else if constexpr (OP == ROW_INV) {
      F::template run<true, Sync>(v, lds, t, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale);
    } else {
      F::template run<false, Sync>(v, lds, t, tw);
    }
    if (row < a.n_rows) {
      float2* p = out_base(row);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int x = t + T * m;
        if (x >= a.store_lo && x < a.store_hi) st<S::NT>(p + XS_OUT * m, v[m]);
      }
    }
  }
}

// Slice step with the transmission row resident: a task is a row group yb
// (slice rows yb * G + g, one per group of T threads) and a chunk of
// image_chunk consecutive images of the probe group, which it runs in turn
// with the slice row held in registers, so a slice row is read once per
// chunk instead of once per image.
template <class S, bool TL = false>
__global__ void __launch_bounds__(kRowThreads, S::W > 0 ? S::W : 1) row_step_tr_kernel(RowArgs a) {
  constexpr int N = S::N, E = S::E;
  using F = Fft<N, E, typename S::RL, S::CT>;
  constexpr int T = F::T;
  constexpr int G = kRowThreads / T;
  using Sync = typename std::conditional<(T <= 64 && !TL), WaveSync, BlockSync>::type;
  using Lds = RowLdsT<S::LS>;
  using Map = RowMap<S, TL>;
  __shared__ float2 smem[G * Lds::padded(N)];

  const int g = Map::g();
  const int t = Map::t();
  Lds lds{smem + g * Lds::padded(N)};
  float2 tw[F::NTW_ALLOC];
  F::twiddles(t, tw);

  const int rows_img = a.rows_per_image;
  const int images = static_cast<int>(a.n_rows / rows_img);
  const int n_groups = (rows_img + G - 1) / G;
  const int chunks = (images + a.image_chunk - 1) / a.image_chunk;
  for (int u = blockIdx.x; u < n_groups * chunks; u += gridDim.x) {
    const int yb = u / chunks, b0 = (u % chunks) * a.image_chunk;
    const int b1 = min(images, b0 + a.image_chunk);
    // a tail group's surplus threads repeat the last row and store nothing
    const bool own = yb * G + g < rows_img;
    const int y = own ? yb * G + g : rows_img - 1;
    float2 tv[E];
    const float2* tr = a.tsl + static_cast<long long>(y) * N + t;
#pragma unroll
    for (int m = 0; m < E; ++m) tv[m] = tr[T * m];
    for (int b = b0; b < b1; ++b) {
      const long long r = static_cast<long long>(b) * rows_img + y;
      constexpr int XS = TL ? 2 * T : T;  // element stride of a row in the wave layout
      static_assert(!TL || T % 4 == 0, "tiled rows need whole tiles per thread group");
      float2* p = TL ? a.psi + tile_offset(r, t, N) : a.psi + r * N + t;
      float2 v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int x = t + T * m;
        v[m] = (x >= a.band_lo && x < a.band_hi) ? ld<S::NT>(p + XS * m) : make_float2(0.f, 0.f);
      }
      F::template run<true, Sync>(v, lds, t, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(v[m], tv[m]);
      F::template run<false, Sync>(v, lds, t, tw);
      if (own) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int x = t + T * m;
          if (x >= a.store_lo && x < a.store_hi) st<S::NT>(p + XS * m, v[m]);
        }
      }
    }
  }
}

// ------------------------------------------------------------------ column pass
template <class S, int OP>
__global__ void __launch_bounds__(S::threads(), S::W > 0 ? S::W : 1) col_kernel(ColArgs a) {
  constexpr int N = S::N, E = S::E, C = S::C;
  using F = Fft<N, E, typename S::RL, S::CT, S::LR>;
  constexpr int T = F::T;
  constexpr int NW = S::threads() / 64;
  constexpr bool LOADS = OP != COL_PROBE;
  using Lds = ColLds<C, S::DB>;
  constexpr int LDS_ONE = Lds::padded(N);
  __shared__ float2 smem[S::DB ? 2 * LDS_ONE : LDS_ONE];
  __shared__ double red[NW > 0 ? NW : 1];
  __shared__ float2 roots[S::LR ? N / 4 : 1];

  const int c = threadIdx.x % C;
  const int jt = threadIdx.x / C;
  Lds lds{smem, c, LDS_ONE};
  if constexpr (S::LR) load_roots<N>(roots, lds, S::threads());
  float2 tw[F::NTW_ALLOC];
  F::twiddles(jt, tw);
  const long long plane = static_cast<long long>(N) * a.Nx;
  const long long stride = a.Nx;

  // Workgroup w owns column strip w % strips for the whole launch and visits
  // images w / strips, w / strips + step, ... (the launcher makes gridDim.x a
  // multiple of strips), so the strip's propagator and probe values are read
  // once into registers; the workgroups resident at one time still cover all
  // strips of a few images, which keeps HBM rows contiguous.
  // XCD-aware virtual workgroup id: the dispatcher deals workgroups
  // round-robin over the XCDs (each with its own L2), so consecutive virtual
  // ids -- adjacent strips of one image, which share 128-byte lines when the
  // strips are narrower than a line -- run on one XCD.
  const int xcd = static_cast<int>(blockIdx.x % kXcds);
  const int q = static_cast<int>(gridDim.x / kXcds), r = static_cast<int>(gridDim.x % kXcds);
  const int vid = xcd * q + min(xcd, r) + static_cast<int>(blockIdx.x / kXcds);
  const int strip = a.strip0 + vid % a.n_strips;
  const int col = strip * C + c;
  const long long step = gridDim.x / a.n_strips;
  // READOUT covers every strip; a strip outside the band has a zero
  // propagator, so its intensities are zero and its wave is never read.
  const bool live = OP != COL_READOUT || !a.apply_pm ||
                    (strip * C < a.band_hi && strip * C + C > a.band_lo);

  // Element (jt + T m, col) of a row-major plane is rm_lane + m T Nx; of the
  // wave in 2 x 4 tiles (a.tiled, T even) st_lane + m T Nx; pm is in either
  // layout (a.pm_tiled).  SB addressing: the byte offsets lane_off (wave) and
  // pm_off plus m * row_block (the launcher checks that a plane's bytes fit
  // 32 bits).
  static_assert(T % 2 == 0, "tiled columns need whole tile rows per element");
  const long long rm_lane = jt * stride + col;
  const long long st_lane = a.tiled ? tile_offset(jt, col, stride) : rm_lane;
  const long long pm_lane = a.tiled && a.pm_tiled ? st_lane : rm_lane;
  const long long m_stride = T * stride;
  const unsigned lane_off = static_cast<unsigned>(st_lane * static_cast<long long>(sizeof(float2)));
  const unsigned pm_off = static_cast<unsigned>(pm_lane * static_cast<long long>(sizeof(float2)));
  const unsigned row_block = static_cast<unsigned>(m_stride * static_cast<long long>(sizeof(float2)));
  const unsigned plane_bytes = static_cast<unsigned>(plane * static_cast<long long>(sizeof(float2)));
  auto pm_at = [&](int m) -> float2 {
    if constexpr (S::SB) return buf_ld<false>(plane_rsrc(a.pm, plane_bytes), pm_off, m * row_block);
    else return a.pm[pm_lane + m * m_stride];
  };

  constexpr bool PM_AT_USE = OP == COL_PROP && S::PG;
  float2 cst[PM_AT_USE ? 1 : E];  // PROP / READOUT: propagator * mask; PROBE: psi_k * (-1)^(row+col)
  if constexpr (PM_AT_USE) {
  } else if constexpr (OP == COL_PROP || OP == COL_READOUT) {
    if (OP == COL_PROP || a.apply_pm) {
      const float sc = OP == COL_PROP ? a.pm_scale : 1.0f;
#pragma unroll
      for (int m = 0; m < E; ++m) cst[m] = cscale(pm_at(m), sc);
    }
  } else if constexpr (OP == COL_PROBE) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int row = jt + T * m;
      const float2 k = a.psi_k[static_cast<long long>(row) * stride + col];
      cst[m] = ((row + col) & 1) ? make_float2(-k.x, -k.y) : k;
    }
  }

  auto load = [&](long long b, float2(&v)[E]) {
    if constexpr (S::SB) {
      const auto img = plane_rsrc(a.psi + b * plane, plane_bytes);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = buf_ld<S::NT>(img, lane_off, m * row_block);
    } else {
      const float2* p = a.psi + b * plane + st_lane;
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = ld<S::NT>(p + m * m_stride);
    }
  };

  float2 pv[E];
  long long b = vid / a.n_strips;
  if constexpr (S::PF && LOADS) {
    if (b < a.B && live) load(b, pv);
  }
  for (; b < a.B; b += step) {
    float2* base = a.psi + b * plane + st_lane;
    float2 v[E];

    if constexpr (OP == COL_PROBE) {
      // psi_k * exp(2 pi i (-kx rx - ky ry)) * (-1)^(row + col)
      // (build_probes_stack + fftshift_2d_stack, probe_kernels.cu:239-304)
      const double ry = a.pos[2 * b];
      const double rx = a.pos[2 * b + 1];
      const float2 ex = unit_phase(-k_lin(col, a.kmax, a.Nx) * rx);
#pragma unroll
      for (int m = 0; m < E; ++m)
        v[m] = cmul(cst[m], cmul(ex, unit_phase(-k_lin(jt + T * m, a.kmax, N) * ry)));
      F::template run<true, BlockSync>(v, lds, jt, tw);
    } else if (!live) {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = make_float2(0.f, 0.f);
    } else {
      if constexpr (S::PF) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = pv[m];
        if (b + step < a.B) load(b + step, pv);
      } else {
        load(b, v);
      }
      if constexpr (OP == COL_PROP) {
        // compact twiddles: rewritten in place once per transform (fewer moves)
        F::template run<false, BlockSync, true>(v, lds, jt, tw);
        if constexpr (PM_AT_USE) {
#pragma unroll
          for (int m = 0; m < E; ++m) v[m] = cmul(v[m], cscale(pm_at(m), a.pm_scale));
        } else {
#pragma unroll
          for (int m = 0; m < E; ++m) v[m] = cmul(v[m], cst[m]);
        }
        F::template run<true, BlockSync, true>(v, lds, jt, tw);
      } else if constexpr (OP == COL_FWD || OP == COL_READOUT) {
        F::template run<false, BlockSync>(v, lds, jt, tw);
      } else if constexpr (OP == COL_INV) {
        F::template run<true, BlockSync>(v, lds, jt, tw);
      }
    }

    if constexpr (OP == COL_READOUT) {
      float inten[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const long long off = (jt + T * m) * stride + col;
        float2 y = v[m];
        if (a.apply_pm) y = cmul(y, cst[m]);
        inten[m] = (y.x * y.x + y.y * y.y) * a.amp;
        a.cbed[b * plane + off] = inten[m];
      }
      if (a.mom_ws) {  // per-strip moments for the record encoder (FP64; the squares are exact)
        double s[2] = {0.0, 0.0};
#pragma unroll
        for (int m = 0; m < E; ++m) {
          s[0] += double(inten[m]);
          s[1] += double(inten[m]) * double(inten[m]);
        }
        for (int k = 0; k < 2; ++k) {
          const double w = wave_sum(s[k]);
          if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
          __syncthreads();
          if (threadIdx.x == 0) {
            double tot = 0.0;
#pragma unroll
            for (int i = 0; i < NW; ++i) tot += red[i];
            a.mom_ws[(b * a.strips_per_image + strip) * 2 + k] = tot;
          }
          __syncthreads();
        }
      }
      for (int d = 0; d < a.n_det; ++d) {
        const float* w = a.det + d * plane;
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < E; ++m) s += double(inten[m]) * double(w[(jt + T * m) * stride + col]);
        s = wave_sum(s);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
          double tot = 0.0;
#pragma unroll
          for (int i = 0; i < NW; ++i) tot += red[i];
          a.det_ws[(b * a.strips_per_image + strip) * a.n_det + d] = tot;
        }
        __syncthreads();
      }
    } else if constexpr (S::SB) {
      const auto img = plane_rsrc(a.psi + b * plane, plane_bytes);
#pragma unroll
      for (int m = 0; m < E; ++m) buf_st<S::NT>(img, lane_off, m * row_block, v[m]);
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) st<S::NT>(base + m * m_stride, v[m]);
    }
  }
}

// ------------------------------------------------------------------ launchers
template <class K>
int persistent_cap(K kernel, int threads) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  return per_cu * device_cu_count();
}

template <class S, int OP, bool TL = false>
int launch_row_op(const RowArgs& a, hipStream_t st) {
  if constexpr (!TL && (OP == ROW_STEP || OP == ROW_INV) && tiled_layout(S::N)) {
    if (a.tiled) return launch_row_op<S, OP, true>(a, st);
  }
  if (!TL && a.tiled) return fail(kErrArg, "row pass: no tiled-layout form of this op at N = %d", S::N);
  if constexpr (S::TR && OP == ROW_STEP) {
    // image chunks sized so that there are about two tasks per resident workgroup
    auto k = row_step_tr_kernel<S, TL>;
    constexpr int G = kRowThreads / (S::N / S::E);
    static const int cap = persistent_cap(k, kRowThreads);
    const int groups = (a.rows_per_image + G - 1) / G;
    const int images = static_cast<int>(a.n_rows / a.rows_per_image);
    const int chunks = std::max(1, std::min(images, (2 * cap + groups - 1) / groups));
    RowArgs c = a;
    c.image_chunk = (images + chunks - 1) / chunks;
    const long long tasks = static_cast<long long>(groups) * ((images + c.image_chunk - 1) / c.image_chunk);
    const int grid = static_cast<int>(std::max<long long>(1, std::min<long long>(tasks, cap)));
    hipLaunchKernelGGL(k, dim3(grid), dim3(kRowThreads), 0, st, c);
    return check_hip(hipGetLastError(), "row pass launch");
  }
  auto k = row_kernel<S, OP, TL>;
  constexpr int G = kRowThreads / (S::N / S::E);
  static const int cap = persistent_cap(k, kRowThreads);
  const long long blocks = (a.n_rows + G - 1) / G;
  const int grid = static_cast<int>(std::max<long long>(1, std::min<long long>(blocks, cap)));
  hipLaunchKernelGGL(k, dim3(grid), dim3(kRowThreads), 0, st, a);
  return check_hip(hipGetLastError(), "row pass launch");
}

template <class S, int OP>
int launch_col_op(ColArgs& a, hipStream_t st) {
  auto k = col_kernel<S, OP>;
  static const int cap = persistent_cap(k, S::threads());
  if (S::SB && static_cast<long long>(S::N) * a.Nx * static_cast<long long>(sizeof(float2)) > 0x7fffffffLL)
    return fail(kErrArg, "column pass: a %d x %d plane exceeds the 32-bit offsets of this schedule", S::N, a.Nx);
  a.strips_per_image = a.Nx / S::C;
  if (OP == COL_PROP) {  // only the strips the propagator band reaches
    a.strip0 = a.band_lo / S::C;
    a.n_strips = (a.band_hi + S::C - 1) / S::C - a.strip0;
  } else {
    a.strip0 = 0;
    a.n_strips = a.strips_per_image;
  }
  if constexpr (S::W > 0) {
    // Schedules sized for several workgroups per CU: one (strip, image) task
    // per workgroup, which the dispatcher deals out to the CU slots as they
    // free up, so the strips x images tasks balance whatever their count.
    const long long grid = static_cast<long long>(a.n_strips) * ((a.B + S::K - 1) / S::K);
    if (grid > 0x7fffffffLL) return fail(kErrArg, "column pass: %lld tasks in one launch", grid);
    hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(grid)), dim3(S::threads()), 0, st, a);
    return check_hip(hipGetLastError(), "column pass launch");
  }
  // Every workgroup keeps one strip for the whole launch, so a launch covers
  // a whole number of strip sets.  With more strips than resident workgroups
  // the strips go in chunks of at most `cap`, largest first; each chunk is
  // its own launch on the same stream.
  const int strip_end = a.strip0 + a.n_strips;
  for (int s0 = a.strip0; s0 < strip_end;) {
    ColArgs c = a;
    c.strip0 = s0;
    c.n_strips = std::min(cap, strip_end - s0);
    const long long sets = std::max<long long>(1, std::min<long long>(c.B, cap / c.n_strips));
    const int grid = static_cast<int>(sets * c.n_strips);
    hipLaunchKernelGGL(k, dim3(grid), dim3(S::threads()), 0, st, c);
    if (int rc = check_hip(hipGetLastError(), "column pass launch")) return rc;
    s0 += c.n_strips;
  }
  return kOk;
}

using RowLaunch = int (*)(int op, const RowArgs&, hipStream_t);
using ColLaunch = int (*)(int op, ColArgs&, hipStream_t);

struct RowSchedEntry {
  const char* name;
  RowLaunch launch;
  const char* precise = nullptr;  // thick-run partner when not derived from the name
};
struct ColSchedEntry {
  const char* name;
  ColLaunch launch;
  int C;
  const char* precise = nullptr;  // thick-run partner when not derived from the name
};
struct SchedTable {
  const RowSchedEntry* rows;
  int n_rows;
  const ColSchedEntry* cols;
  int n_cols;
};

// The base schedule (entry 0 of a table) serves every op; the alternatives
// only replace the hot ops (row STEP, column PROP and READOUT) and defer the
// rest to the base schedule.
template <class S, class Base>
int launch_row(int op, const RowArgs& a, hipStream_t st) {
  if (op == ROW_STEP) return launch_row_op<S, ROW_STEP>(a, st);
  if constexpr (std::is_same<S, Base>::value) {
    if (op == ROW_INV) return launch_row_op<S, ROW_INV>(a, st);
    return launch_row_op<S, ROW_FWD>(a, st);
  } else {
    return launch_row<Base, Base>(op, a, st);
  }
}

template <class S, class Base>
int launch_col(int op, ColArgs& a, hipStream_t st) {
  if (op == COL_PROP) return launch_col_op<S, COL_PROP>(a, st);
  // the multi-workgroup schedules' readout (one task per (strip, image),
  // detector reductions per task) is 20-50% slower than the base schedule's
  if (op == COL_READOUT) {
    if constexpr (S::W > 0)
      return launch_col<Base, Base>(op, a, st);
    else
      return launch_col_op<S, COL_READOUT>(a, st);
  }
  if constexpr (std::is_same<S, Base>::value) {
    if (op == COL_PROBE) return launch_col_op<S, COL_PROBE>(a, st);
    if (op == COL_FWD) return launch_col_op<S, COL_FWD>(a, st);
    return launch_col_op<S, COL_INV>(a, st);
  } else {
    return launch_col<Base, Base>(op, a, st);
  }
}

template <int N>
const SchedTable* sched_table();

}  // namespace namsa
