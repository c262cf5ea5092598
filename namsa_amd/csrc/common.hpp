// Shared helpers for the namsa_amd HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace namsa {

// Return codes of every extern "C" entry point.
enum Status : int {
  kOk = 0,
  kErrArg = -1,          // invalid argument (size, pointer, count)
  kErrUnsupported = -2,  // grid size without a compiled FFT schedule
  kErrHip = -3,          // a HIP runtime call failed
};

// Record a message for namsa_last_error() (thread-local) and return `code`.
int fail(int code, const char* fmt, ...);
// Map a hipError_t to kErrHip with a message; kOk when err == hipSuccess.
int check_hip(hipError_t err, const char* what);
// Number of compute units of the current device (cached per device).
int device_cu_count();
// Look up the current device's address of the fixed scratch array of
// namsa_multislice_noalloc (once per device; namsa_init calls it).
int scratch_setup();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------------ complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Linspace k coordinate of the reference k-grid: k_j = j*kmax/(n-1) - kmax/2,
// evaluated in double exactly as calc_krad does (probe_kernels.cu:11-16).
__device__ __forceinline__ double k_lin(int j, double kmax, int n) {
  return double(j) * kmax / double(n - 1) - kmax / 2.0;
}

}  // namespace namsa
