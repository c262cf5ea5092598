/*
 * namsa_hip.h -- C ABI of the MI355X (gfx950) multislice library libnamsa_hip.so.
 *
 * Every entry point replaces a piece of the reference nlaanait/namsa MSAGPU
 * path, which launched pycuda kernels (namsa/cuda_kernels.py:8-56) and
 * scikit-cuda cuFFT plans from Python.  Conventions:
 *   - all array pointers are DEVICE pointers owned by the caller (e.g. torch
 *     tensors' data_ptr()); complex arrays are interleaved complex64
 *     (float re, float im), C order;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream); every
 *     call only enqueues work on it and never synchronises the host;
 *   - no entry point allocates device memory; scratch space is passed in
 *     (namsa_multislice_noalloc uses a fixed per-device array instead);
 *   - every call returns 0 on success or a negative status; the message of
 *     the last failure on the calling thread is namsa_last_error();
 *   - grid sizes Ny, Nx must be supported FFT lengths (namsa_fft_supported).
 */
#ifndef NAMSA_HIP_H
#define NAMSA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NAMSA_ABI_VERSION 3

#define NAMSA_OK 0
#define NAMSA_ERR_ARG (-1)
#define NAMSA_ERR_UNSUPPORTED (-2)
#define NAMSA_ERR_HIP (-3)

/* Output selector of namsa_multislice. */
#define NAMSA_OUT_CBED 0     /* transmit=True: float32 CBED intensities   */
#define NAMSA_OUT_INCIDENT 1 /* transmit=False: complex incident probes   */
#define NAMSA_OUT_EXIT 2     /* complex real-space exit waves              */
/* Or-ed into namsa_multislice's mode with NAMSA_OUT_CBED or NAMSA_OUT_EXIT
 * where namsa_wave_tiled(Ny, Nx): pm is given in the 2 x 4 tile layout
 * (namsa_tile_plane), which the tiled column pass then reads in 64-byte
 * pieces too.  No counterpart in the reference. */
#define NAMSA_PM_TILED 0x100

int namsa_abi_version(void);
const char* namsa_last_error(void);

/* Select the device (replaces setup_device, namsa/msa.py:535-553). */
int namsa_init(int device);

/* 1 if n is a compiled transform length, else 0. */
int namsa_fft_supported(int n);

/*
 * Compiled schedules of the two passes of the slice step, per transform
 * length (pass 0 = row pass, 1 = column pass).  No reference counterpart: the
 * reference specialises its kernels by Jinja2 constants and JIT-compiles them
 * per run (namsa/cuda_kernels.py:8-56); here every length has several
 * precompiled schedules and the autotuner (namsa_amd/autotune.py) picks one.
 * The selection is process-wide; change it between launches, not during them.
 *   namsa_sched_count  number of schedules (0 for an unsupported length)
 *   namsa_sched_name   name of schedule v, or NULL
 *   namsa_sched_get    selected schedule, or a negative status
 *   namsa_sched_set    select schedule v
 */
int namsa_sched_count(int n, int pass);
const char* namsa_sched_name(int n, int pass, int v);
int namsa_sched_get(int n, int pass);
int namsa_sched_set(int n, int pass, int v);
/* Compact-twiddle (".ct") schedules keep one register set of twiddles per
 * butterfly and form the rest; namsa_multislice runs for more than
 * namsa_precise_slices(n) slices use namsa_sched_precise(n, pass, v) -- the
 * schedule holding every twiddle that v derives from (v itself otherwise). */
int namsa_precise_slices(int n);
int namsa_sched_precise(int n, int pass, int v);

/*
 * Kirkland projected-potential tables, one Py x Px float32 table per element
 * (replaces MSA.calc_atomic_potentials / _get_potential, namsa/msa.py:137-146,
 * :180-196, and bin_2d_array, namsa/utils.py:101-108).
 *   params:  [nZ][4][3] float64 rows a, b, c, d (kirkland_params[Z-1])
 *   tables:  [nZ][Py][Px] float32 output
 * Subsamples sit on linspace(-range/2, range/2, oversample*P) per axis, the
 * table is their oversample x oversample block mean shifted so its minimum is 0.
 */
int namsa_atomic_tables(const double* params, int nZ, int Py, int Px, double range,
                        int oversample, float* tables, void* stream);

/*
 * Transmission slices t = exp(i sigma V) (replaces BuildScatteringPotential,
 * namsa/potential_kernels.cu:5-60, launched at namsa/msa.py:626-639).
 *   tables:        [nZ][Py][Px] float32
 *   sites:         [A][3] int32 (table index, y pixel, x pixel), grouped by slice
 *   slice_offsets: [S+1] int32, sites of slice s are [offsets[s], offsets[s+1])
 *   t_out:         [S][Ny][Nx] complex64 output
 * Site footprint rows are rintf(y - Py/2) <= row < rintf(y + Py/2), clipped to
 * the grid and to the table; sites with a negative table index are skipped.
 * NaN components of t are replaced by 1.
 */
int namsa_build_slices(const float* tables, int nZ, int Py, int Px, const int* sites,
                       const int* slice_offsets, int S, int Ny, int Nx, float sigma,
                       void* t_out, void* stream);

/*
 * Projected potential V[S][Ny][Nx] float32 of the same site lists (the
 * accumulation of namsa_build_slices without the exp(i sigma V) step); the
 * corrector network takes the projected potential as an input channel.
 */
int namsa_build_potential(const float* tables, int nZ, int Py, int Px, const int* sites,
                          const int* slice_offsets, int S, int Ny, int Nx, float* v_out,
                          void* stream);

/*
 * Slices of a sampled 3D potential, e.g. from an all-electron ab initio
 * calculation (replaces the trapezoid integration of
 * MSA.integrate_potential_slices, namsa/msa.py:373-424, and the
 * exp(1j*sigma*V) of MSAGPU.integrate_potential_slices, msa.py:568-571).
 *   pot:   [S*d][Ny][Nx] float64 samples, z-major (the caller crops z to S*d
 *          samples and the plane to the grid, as msa.py:386-397 does)
 *   z:     [S*d] float64 sample coordinates
 *   t_out: [S][Ny][Nx] complex64 exp(i sigma V_s), or NULL
 *   v_out: [S][Ny][Nx] float64 V_s, or NULL (at least one of the two)
 * V_s = -trapz(pot[s*d .. s*d+d-1], z[s*d .. s*d+d-1]) over its own d samples;
 * no interval spans two slices, as with the reference's np.split + trapz.
 */
int namsa_integrate_slices(const double* pot, const double* z, int S, int d, int Ny, int Nx, double sigma,
                           void* t_out, double* v_out, void* stream);

/*
 * Probe in k-space psi_k = A(k) exp(-i chi(k)) on the linspace k-grid
 * (replaces soft_aperture / hard_aperture / spherical_phase_error /
 * mult_wise_c2d_re2d, namsa/probe_kernels.cu:180-222 and :126-134, launched at
 * namsa/msa.py:711-715).
 *   smooth != 0: A = 1/(1+exp(-2*steepness*(k_semi-k))); else A = [k < k_semi].
 *   psi_k: [Ny][Nx] complex64 output; apert: [Ny][Nx] float32 output or NULL.
 */
int namsa_probe_k(int Ny, int Nx, float kmax, float k_semi, float lambda, float C1, float C3,
                  float C5, int smooth, float steepness, void* psi_k, float* apert,
                  void* stream);

/*
 * Fresnel propagator times bandwidth mask (replaces propagator, hard_aperture
 * and mult_wise_c2d_re2d, namsa/probe_kernels.cu:167-194, at
 * namsa/msa.py:833-839).
 *   pm:   [Ny][Nx] complex64 output exp(-i pi k^2 lambda dz) * [k < k_cut]
 *   mask: [Ny][Nx] float32 output or NULL
 */
int namsa_propagator(int Ny, int Nx, float kmax, float k_cut, float lambda, float dz,
                     void* pm, float* mask, void* stream);

/*
 * Batched unnormalised 2D C2C transform in place (replaces the scikit-cuda
 * cuFFT plan + cufftExecC2C calls, namsa/msa.py:721-722, :880, :962, :976,
 * :981, :987).  inverse = 0: exp(-2 pi i ...), inverse = 1: exp(+2 pi i ...).
 *   data: [B][Ny][Nx] complex64
 */
int namsa_fft2d(void* data, int B, int Ny, int Nx, int inverse, void* stream);

/*
 * Thin-object approximation (replaces namsa/utils.py:130-135,
 * thin_object_approx): out = |FFT2(probes * exp(i sum_s V_s))|, unshifted and
 * unnormalised, the phase summed over slices in FP64.
 *   V:      [S][Ny][Nx] float64
 *   probes: [B][Ny][Nx] complex64
 *   psi_ws: [B][Ny][Nx] complex64 workspace
 *   out:    [B][Ny][Nx] float32
 */
int namsa_thin_object(const double* V, int S, const void* probes, int B, int Ny, int Nx, void* psi_ws, float* out,
                      void* stream);

/* sum |x|^2 over n complex64 values into *out (one float64 on the device);
 * the normalization of build_probe, namsa/msa.py:727. */
int namsa_norm2(const void* data, long long n, double* out, void* stream);

/*
 * Detector integrals of stored CBEDs (replaces the host sum of
 * MSA.integrate_cbed, namsa/msa.py:436-439).
 *   cbed: [B][Ny][Nx] float32; det: [nDet][Ny][Nx] float32;
 *   out:  [B][nDet] float64 = sum over pixels of cbed * det, in float64.
 */
int namsa_integrate(const float* cbed, int B, int Ny, int Nx, const float* det, int nDet,
                    double* out, void* stream);

/*
 * Number of float64 scratch values namsa_multislice needs for the detector
 * partial sums of B probes and nDet detectors.
 */
size_t namsa_det_workspace_doubles(int B, int Ny, int Nx, int nDet);

/*
 * Building blocks of namsa_multislice, one launch each, for callers that
 * schedule slices themselves (graph capture, per-pass timing).  The wave
 * state between them is internal: after namsa_probe_init and after every
 * namsa_slice_cols it holds C^-1 of the spectrum of the current wave; after
 * namsa_slice_rows it holds the row transform of t_i * psi_i.
 *   namsa_probe_init  probe stack (build_probes_stack + IFFT + fftshift_2d_stack,
 *                     namsa/probe_kernels.cu:239-304, msa.py:959-964)
 *   namsa_slice_rows  first half of one slice (msa.py:973-976): row pass
 *   namsa_slice_cols  second half of one slice (msa.py:976-981): column pass
 * A slice i is namsa_slice_rows(t + i*Ny*Nx) followed by namsa_slice_cols.
 * band_lo, band_hi: the columns of pm that may be nonzero (see below); the
 * row pass right after namsa_probe_init must use the full band [0, Nx).
 * namsa_slice_rows reads only the band and stores only the band widened to
 * multiples of 16 columns, the only columns the namsa_slice_cols with the
 * same band reads (with the full band it stores every column).
 */
int namsa_probe_init(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                     void* psi_ws, void* stream);
int namsa_slice_rows(void* psi_ws, int B, int Ny, int Nx, const void* t_slice, int band_lo,
                     int band_hi, void* stream);
int namsa_slice_cols(void* psi_ws, int B, int Ny, int Nx, const void* pm, int band_lo, int band_hi,
                     void* stream);

/*
 * The same two passes over a wave stored in 2 x 4 tiles: element (y, x) of a
 * [Ny][Nx] plane at ((y/2)(Nx/4) + x/4)*8 + (y%2)*4 + x%4 (the column pass
 * takes pm in the same tiles; t_slice stays row-major).  In CBED and
 * exit-wave mode namsa_multislice keeps its wave in this layout at the sizes
 * where namsa_wave_tiled(Ny, Nx) returns 1 (2048 x 2048): the column pass
 * then reads 64-byte pieces instead of 32-byte row segments, and in
 * exit-wave mode the last pass writes the exit waves back row-major in
 * place.  The tiled entry points return NAMSA_ERR_UNSUPPORTED at other
 * sizes.  No counterpart in the
 * reference (its cuFFT plans own their layouts); for per-pass timing and
 * custom slice loops at those sizes.
 */
int namsa_wave_tiled(int Ny, int Nx);
int namsa_slice_rows_tiled(void* psi_ws, int B, int Ny, int Nx, const void* t_slice, int band_lo,
                           int band_hi, void* stream);
int namsa_slice_cols_tiled(void* psi_ws, int B, int Ny, int Nx, const void* pm, int band_lo, int band_hi,
                           void* stream);
/* dst[tile_offset(y, x)] = src[y][x]: a [Ny][Nx] complex64 plane into the
 * 2 x 4 tile layout (Ny even, Nx a multiple of 4; dst != src).  For the
 * propagator of NAMSA_PM_TILED. */
int namsa_tile_plane(const void* src, void* dst, int Ny, int Nx, void* stream);

/*
 * Batched multislice, fused (replaces MSAGPU.__propagate_beams,
 * namsa/msa.py:926-992: build_probes_stack, fftshift_2d_stack,
 * mult_wise_c3d_c3d_ind, mult_wise_c3d_c2d, mod_square_stack and the four
 * cuFFT executions per slice; and the detector sums of integrate_cbed,
 * namsa/msa.py:426-439).
 *   psi_k:   [Ny][Nx] complex64 probe in k-space
 *   pos_yx:  [B][2] float32 probe positions (y, x) in Angstrom
 *   t:       [S][Ny][Nx] complex64 transmission slices, applied in ascending order
 *   pm:      [Ny][Nx] complex64 propagator * mask
 *   band_lo, band_hi: every column of pm outside [band_lo, band_hi) is zero
 *            (the bandwidth mask); those columns are never computed.  (0, 0)
 *            means the full width.
 *   group:   probes pushed through all slices together (0 = all B); a group
 *            whose waves fit the 256 MB last-level cache keeps the slice
 *            loop out of HBM
 *   mode:    NAMSA_OUT_CBED     cbed[B][Ny][Nx] = |sqrt(N^2) FFT(psi_S/N^2)|^2 * out_scale
 *                               (out_scale = 1/normalization), plus
 *                               det_out[B][nDet] = sum(cbed * det[d]) if nDet > 0
 *            NAMSA_OUT_INCIDENT psi_ws[B][Ny][Nx] = incident probes * out_scale
 *            NAMSA_OUT_EXIT     psi_ws[B][Ny][Nx] = real-space exit waves * out_scale
 *   psi_ws:  complex64 wave state: min(B, group) planes in NAMSA_OUT_CBED mode
 *            (every group reuses them), B planes (the output) otherwise
 *   det:     [nDet][Ny][Nx] float32 detector weights (may be NULL if nDet == 0)
 *   det_ws:  namsa_det_workspace_doubles(B, Ny, Nx, nDet) float64 scratch values
 *   det_out: [B][nDet] float64
 * The passes run the schedules namsa_sched_set selected; for S > 300 (S > 200
 * for lengths divisible by 3) the slice loop and the readout use the schedules
 * holding every twiddle in place of compact-twiddle (".ct") ones, whose
 * round-off would approach 1e-4 of the CBED by 1000 slices.
 */
int namsa_multislice(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                     const void* t, int S, const void* pm, int band_lo, int band_hi, int group,
                     int mode, float out_scale, void* psi_ws, float* cbed, const float* det,
                     int nDet, double* det_ws, double* det_out, void* stream);

/*
 * namsa_multislice in CBED mode that also returns, per probe, the sums of
 * the CBED and of its square (FP64, taken in the readout from the values it
 * writes). They are the statistics process_cbed needs
 * (scripts/summit_scripts/utils.py:104-118), so that
 * namsa_cbed_records_from_moments encodes the records in one pass.
 *   mom_ws:  namsa_det_workspace_doubles(group or B, Ny, Nx, 2) doubles
 *   mom_out: [B][2] doubles (sum, sum of squares)
 */
int namsa_multislice_moments(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                             const void* t, int S, const void* pm, int band_lo, int band_hi, int group,
                             float out_scale, void* psi_ws, float* cbed, const float* det, int nDet,
                             double* det_ws, double* det_out, double* mom_ws, double* mom_out, void* stream);

/*
 * namsa_multislice in NAMSA_OUT_CBED mode that allocates nothing and takes no
 * workspace: the wave state of one probe group and its detector partial sums
 * live in a fixed device array of NAMSA_SCRATCH_BYTES that each GPU the
 * library runs on holds until the process ends.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops with a fixed memory budget).
 *   group: probes per group; 0 = as many as the scratch holds, i.e.
 *          floor(NAMSA_SCRATCH_BYTES / (8 Ny Nx + 8 (Nx/4) nDet)) (31 at
 *          512^2 with two detectors, 1 at 2048^2).  A larger explicit group,
 *          or a grid of which one probe does not fit, is rejected
 *          (NAMSA_ERR_ARG) before anything is enqueued.
 * Other arguments and results as namsa_multislice.  Every call on one GPU
 * shares that GPU's array, and calls that overlap on one GPU (two streams,
 * two threads) race; the array keeps the last call's contents, none of which
 * a call reads before writing it.  namsa_init looks up the array's address
 * for its device; the first call on a device that skipped namsa_init does so.
 */
#define NAMSA_SCRATCH_BYTES (64u << 20)
int namsa_multislice_noalloc(const void* psi_k, const float* pos_yx, int B, int Ny, int Nx, float kmax,
                             const void* t, int S, const void* pm, int band_lo, int band_hi, int group,
                             float out_scale, float* cbed, const float* det, int nDet, double* det_out,
                             void* stream);

/*
 * Row transforms of planes in place: forward (unscaled, scale must be 1) or
 * inverse times scale.  The slice state between steps is the row transform
 * of the wave, so the inverse with scale 1/Nx gives the wave in real space
 * and the forward takes a real-space wave back to the state.  No counterpart
 * in the reference (its state is the real-space wave, msa.py:971-981); the
 * coarse-slice corrector uses it to read and correct the wave.
 */
int namsa_row_fft(void* data, int B, int Ny, int Nx, int inverse, float scale, void* stream);

/*
 * ---- Coarse-slice corrector network (no counterpart in the reference) ----
 * Strided batched GEMM on the matrix cores, BF16 operands, FP32 accumulate:
 *   C[b](i,j) = alpha sum_k A[b](i,k) B[b](k,j) + beta C[b](i,j) (+ bias, GELU)
 * A(i,k) at A + b*sAb + i*sAi + k*sAk (FP32), likewise B(k,j) and C(i,j).
 * split_k > 1 splits K over workgroups and adds into C with FP32 atomics
 * (bias and GELU must then be off).  bias_mode 1: + bias[i], 2: + bias[j].
 * gelu: C = gelu(value) and P (same strides as C) = value.
 */
int namsa_gemm_bf16(const float* A, long long sAb, long long sAi, long long sAk, const float* B,
                    long long sBb, long long sBk, long long sBj, float* C, long long sCb,
                    long long sCi, long long sCj, int M, int N, int K, int batch, int split_k,
                    float alpha, float beta, const float* bias, int bias_mode, int gelu, float* P,
                    void* stream);

/* x[0..n) = v */
int namsa_fill(float* x, long long n, float v, void* stream);
/* g *= gelu'(pre), tanh form of GELU */
int namsa_gelu_backward(float* g, const float* pre, long long n, void* stream);
/* Adam with decoupled weight decay; step counts from 1 */
int namsa_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
               float eps, int step, float wd, void* stream);
/*
 * Mode bookkeeping of the spectral layers (m retained modes per axis, C
 * channels, B samples).  combine: per (c,b) the 2m x 2m product
 * P[(r,p)][(t,q)] of [cos; sin](y, p) with the x-transformed rows
 * Z[y][(t,q)] = (sum x cos, -sum x sin) becomes
 * Xs[p*m+q][(0,c)][b] = P[0p][0q] + P[1p][1q] (real part) and
 * Xs[p*m+q][(1,c)][b] = P[0p][1q] - P[1p][0q] (imaginary part).
 * expand: Ys[p*m+q][(r,o)][b] becomes, per (o,b), the 2m x 2m operand
 * E[(r,p)][(t,q)] with E[0p][0q] = Yr, E[1p][0q] = -Yi, E[0p][1q] = Yi,
 * E[1p][1q] = Yr, so that [cos | sin](y, (r,p)) E gives the y-inverse
 * transform (real, imaginary) in columns (t,q).
 */
int namsa_modes_combine(const float* P, float* Xs, int C, int B, int m, void* stream);
int namsa_modes_expand(const float* Ys, float* E, int C, int B, int m, void* stream);
/*
 * Channel-last BF16 inference of the corrector network (32 channels, 2m = 32
 * row-DFT columns, head width 64), no counterpart in the reference.
 * Activations x: bf16 [B*H][W][32]; Z / Zp: float [(c*B + b)*H + y][32] (the
 * row DFT and row-inverse operands the strided GEMM path's mode-space
 * products read and write); fx: bf16 [Wp][32] = Fx[x][q] (the truncated row
 * DFT, rows >= W zero), fxT: bf16 [32][Wp] its transpose (columns >= W zero),
 * Wp a multiple of 32 >= W.
 *   namsa_fno_lift:  x0 = Wl u + bl (u float [cin][B*H*W], cin <= 8, Wl
 *                    [32][cin]) and Z = row DFT of x0
 *   namsa_fno_layer: x_out = gelu(alpha Fx Zp + Wt x_in + bias) and Z_next =
 *                    row DFT of x_out; with out != NULL (the last layer) no
 *                    GELU and out[t][p] = Q2[t] . gelu(Q1 x + bq1) + bq2[t]
 *                    (float [2][B*H*W]) instead of x_out and Z_next.
 */
int namsa_fno_lift(const float* u, int cin, const float* Wl, const float* bl, const void* fxT, int B, int H, int W,
                   int Wp, void* x0, float* Z, void* stream);
int namsa_fno_layer(const void* x_in, const float* Zp, const void* fx, const void* fxT, const float* Wt,
                    const float* bias, float alpha, int B, int H, int W, int Wp, void* x_out, float* Z_next,
                    const float* Q1, const float* bq1, const float* Q2, const float* bq2, float* out, void* stream);

/*
 * Spectral (one-shot) corrector: works on the centred exit spectrum Psi(k) of
 * the coarse path, |Psi|^2 being the CBED, inside the window
 * [r0, r0+H) x [c0, c0+W) of the Ny x Nx plane that bounds the bandwidth mask.
 * psi_c, psi_0, target, cbed: full planes [B][Ny][Nx]; u, out, gout:
 * channel-major [ch][B*H*W] over the window.
 *
 * namsa_spectral_pack: network inputs
 *   u[0..1] = s Psi_c (coarse exit spectrum), u[2..3] = s Psi_0 (the probe's
 *   exit spectrum through vacuum of the same thickness), u[4] = phase_scale *
 *   |k|^2 on sample b's linspace k-grid of extent kmax[b] (device [B]),
 *   u[5] = thick[b] * thick_scale (device [B]), then with
 *   opts & 1: log1p(vscale |vhat[vidx[b]]|) (vhat: [nv][Ny][Nx] complex,
 *   centred spectra of the projected potential; vidx NULL = plane 0; an
 *   index outside [0, nv) gives NaN, never a read outside vhat), with
 *   opts & 2: |s Psi_c|.
 * namsa_spectral_loss: with Psi = Psi_c (1 + o), o = out[0] + i out[1],
 *   I = |s Psi|^2 and F = s^2 target: sums[2b] = sum (I - F)^2,
 *   sums[2b+1] = sum F^2 over the window; gout (optional) = gradient of
 *   mean_b sums[2b] / sums[2b+1] with respect to out.
 * namsa_spectral_apply: cbed = |Psi_c (1 + o)|^2 * scale in the window,
 *   |Psi_c|^2 * scale elsewhere.
 * namsa_real_to_complex: out = (scale v, 0), times (-1)^(y+x) if checker (so
 *   that its forward transform comes out centred), over [B][Ny][Nx].
 */
int namsa_spectral_pack(const void* psi_c, const void* psi_0, const void* vhat, const int* vidx, int nv, int B, int Ny,
                        int Nx, int r0, int c0, int H, int W, const float* kmax, float s, float phase_scale,
                        const float* thick, float thick_scale, float vscale, int opts, float* u, void* stream);
int namsa_spectral_loss(const float* out, const void* psi_c, const float* target, int B, int Ny, int Nx, int r0,
                        int c0, int H, int W, float s, double* sums, float* gout, void* stream);
int namsa_spectral_apply(const float* out, const void* psi_c, int B, int Ny, int Nx, int r0, int c0, int H, int W,
                         float scale, float* cbed, void* stream);
int namsa_real_to_complex(const float* v, int B, int Ny, int Nx, float scale, int checker, void* out,
                          void* stream);

/* ---- training records (the 4D-STEM dataset step after the slice loop) ---- */

/*
 * Per-CBED standardisation to IEEE half (replaces process_cbed,
 * scripts/summit_scripts/utils.py:104-118): out = (x - mean) / std per
 * [Ny][Nx] item, population std, float32 arithmetic for the subtraction and
 * division, round-to-nearest-even to half. A constant item gives NaN, as
 * numpy's 0/0 does.
 *   cbed: [P][Ny][Nx] float32;  out_f16: [P][Ny][Nx] half (uint16 bits)
 *   ws:   namsa_records_workspace_doubles(P, Ny, Nx) doubles
 */
long long namsa_records_workspace_doubles(int P, int Ny, int Nx);
int namsa_cbed_records(const float* cbed, int P, int Ny, int Nx, void* out_f16, double* ws, void* stream);
/* The same records from the moments namsa_multislice_moments returned: one pass. */
int namsa_cbed_records_from_moments(const float* cbed, const double* moments, int P, int Ny, int Nx,
                                    void* out_f16, void* stream);

/*
 * Projected-potential training record (replaces process_potential,
 * scripts/summit_scripts/utils.py:74-102): mean over slices of imag(t),
 * scipy.ndimage.gaussian_filter(sigma, truncate, mode 'reflect'), the crop
 * rows and columns [Ny/4, min(3Nx/4, dim)), standardised with the deviation
 * floored at 1/sqrt(size), to half.
 *   t:       [S][Ny][Nx] complex64 transmission slices
 *   out_f16: [rows][cols] half, rows/cols from namsa_potential_record_shape
 *   ws:      namsa_potential_record_workspace_bytes(Ny, Nx) bytes
 * The reference uses sigma = 1.2 and scipy's default truncate = 4.0.
 */
int namsa_potential_record_shape(int Ny, int Nx, int* rows, int* cols);
long long namsa_potential_record_workspace_bytes(int Ny, int Nx);
int namsa_potential_record(const void* t, int S, int Ny, int Nx, double sigma, double truncate, void* out_f16,
                           void* ws, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NAMSA_HIP_H */
