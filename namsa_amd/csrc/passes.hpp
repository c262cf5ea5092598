// Row and column passes of the fused slice step, shared by the ABI drivers.
#pragma once
#include "common.hpp"

namespace namsa {

enum RowOp { ROW_STEP = 0, ROW_INV = 1, ROW_FWD = 2 };
enum ColOp { COL_PROP = 0, COL_PROBE = 1, COL_READOUT = 2, COL_FWD = 3, COL_INV = 4 };
enum Pass { PASS_ROW = 0, PASS_COL = 1 };

// Row pass over rows of length N = Nx: psi [n_rows][N], tsl [rows_per_image][N].
// Entries outside [band_lo, band_hi) are known to be zero and are not loaded.
struct RowArgs {
  float2* psi;
  long long n_rows;
  int rows_per_image;
  const float2* tsl;  // STEP: transmission slice
  float scale;        // INV: output scale
  int band_lo, band_hi;    // columns loaded (the rest are zero on input)
  int store_lo, store_hi;  // columns stored (the rest are never read afterwards)
  int image_chunk;         // STEP with a resident transmission row: images per task (set by the launcher)
  int tiled;               // STEP: the wave is stored in 2 x 4 tiles (tile_offset), not row-major
};

// Column pass over columns of length N = Ny of B images [B][Ny][Nx].  The
// launcher covers strips [strip0, strip0 + n_strips) of width C; READOUT and
// the transforms cover every strip, PROP only those meeting [band_lo, band_hi)
// (outside that band the propagator, and so the result, is zero).
struct ColArgs {
  float2* psi;
  int B;
  int Nx;
  int strips_per_image;  // Nx / C, filled in by the launcher
  int strip0, n_strips;  // filled in by the launcher
  int band_lo, band_hi;  // columns where pm may be nonzero
  const float2* pm;      // [Ny][Nx] propagator * mask (PROP, READOUT)
  float pm_scale;        // PROP: factor folded into pm (1/N^2)
  const float2* psi_k;   // [Ny][Nx] (PROBE)
  const float* pos;      // [B][2] (y, x) (PROBE)
  double kmax;           // (PROBE)
  float amp;             // intensity factor (READOUT)
  int apply_pm;          // (READOUT)
  float* cbed;           // [B][Ny][Nx] (READOUT)
  const float* det;      // [n_det][Ny][Nx] (READOUT)
  int n_det;
  double* det_ws;        // [B][strips_per_image][n_det] (READOUT)
  double* mom_ws;        // [B][strips_per_image][2]: sums of the CBED and of its square (READOUT), or null
  int tiled;             // PROBE, PROP, READOUT: the wave is stored in 2 x 4 tiles (tile_offset); psi_k,
                         // cbed and det stay row-major
  int pm_tiled;          // with tiled: pm is in the same tiles
};

// Tiled wave layout: element (y, x) of a [Ny][Nx] plane sits in the 64-byte
// tile of rows 2(y/2), 2(y/2)+1 and columns 4(x/4) .. 4(x/4)+3, the tiles in
// row-major order.  A row pass (two rows per workgroup) then reads and writes
// whole tiles, a column pass over 4-column strips 64-byte pieces instead of
// the 32-byte row segments of the row-major layout.  For rows y = j + T m
// with T even and columns x = t + T m with T a multiple of 4 the offset stays
// a lane base plus m times a constant (T Nx and 2 T).
__host__ __device__ constexpr long long tile_offset(long long y, long long x, long long Nx) {
  return (y >> 1) * 2 * Nx + (x >> 2) * 8 + (y & 1) * 4 + (x & 3);
}
// Sizes whose slice loop keeps the wave tiled.  At 2048^2 (6 images, 4-column
// strips) the memory accesses of a column pass alone take 86 us row-major
// and 62 us tiled, those of a row pass 39 and 49 us; with the transforms the
// slice step drops from 165 to 135 us.
constexpr bool tiled_layout(int n) { return n == 2048; }

// MI355X has 8 XCDs, each with its own L2; the dispatcher deals workgroups
// round-robin over them in workgroup-id order.
constexpr int kXcds = 8;

// Narrowest column tile of any compiled schedule: detector scratch is sized
// for it so that any schedule choice fits.
constexpr int kMinColTile = 4;
constexpr int kMaxColTile = 16;
// row steps followed by a band-limited column pass store this-aligned bands
constexpr int kStoreAlign = kMaxColTile;
// Thicker runs than this take the slice loop's passes from the schedules that
// hold every twiddle instead of compact-twiddle ones.  With both passes
// compact the CBED drifts by about 9e-8 per slice at the power-of-two
// lengths and 1.3e-7-1.6e-7 at the lengths with a factor 3 (radix-12 first
// stages; the full forms 4.5e-8 and 7e-8), so the switch keeps every
// thickness below ~6e-5.
constexpr int precise_slices(int n) { return n % 3 == 0 ? 200 : 300; }

bool supported(int n);
int row_pass(int Nx, int op, const RowArgs& a, hipStream_t st, bool precise = false);
// Launches the selected column schedule; on return a.strips_per_image holds
// the strip count that was used.
int col_pass(int Ny, int op, ColArgs& a, hipStream_t st, bool precise = false);

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

}  // namespace namsa
