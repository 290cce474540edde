// Internal declarations shared by the HIP translation units of libgclsph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace gclsph {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes returned through the C ABI.
enum : int {
  kOk = 0,
  kErrArgument = -1,   // bad argument / dimension / length (reference: error(...))
  kErrHip = -2,        // HIP runtime failure
  kErrState = -3,      // call not valid in the object's state (e.g. pair view past capacity)
  kErrNoDevice = -4,
  kErrUnsupported = -5,
};

#define GCL_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      throw ::gclsph::Error(::gclsph::kErrHip, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define GCL_CHECK(cond, code, msg)                                 \
  do {                                                             \
    if (!(cond)) throw ::gclsph::Error((code), std::string(msg));  \
  } while (0)

// Message returned by gclsph_last_error() on this thread (api.hip).
void set_last_error(const std::string& m);

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return unsigned(g < 1 ? 1 : g);
}

// Per-device control block of an SPH problem.  Lives in device memory so a whole
// batch of steps runs without a host round trip; the host reads it once at the
// end of a `stepsolve` call.
struct StepCtl {
  int parity;         // which of the ping-pong state buffers is current
  int rebuild;        // rebuild the cell order at the start of the next step
  int rebuilt;        // the current step rebuilt (membership uses the current x)
  int synced;         // half-step state equals the main state (no correction ran)
  int pparity;       // which ptype / id buffer is current (flipped by reorders only)
  int xrb_cur;       // rebuild positions are the current positions (no step since)
  int iter;          // 1-based iteration inside the running stepsolve call
  int nsteps;        // iterations of the running call
  long long updaten;  // rebuilds in this call
  long long mppcell;  // reference-visible max-particles-per-cell bookkeeping
  double dt;
  double etime;
  double maxcdx_out;  // max over steps of max |cdx|
  // reduction slots (non-negative doubles compared as uint64 bit patterns)
  unsigned long long red_cdx;
  unsigned long long red_acc;
  unsigned long long red_visc;
  unsigned long long red_maxcell;
};

// Device helpers ------------------------------------------------------------
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w > v || w != w) ? w : v;
  }
  return v;
}

// atomic max of a non-negative double (NaN propagates: its bit pattern is larger
// than +Inf's).
__device__ inline void atomic_max_nonneg(unsigned long long* slot, double v) {
  unsigned long long b = __double_as_longlong(v);
  if (b & 0x8000000000000000ull) b = 0;  // -0.0 and negatives clamp to 0
  atomicMax(slot, b);
}

__device__ inline double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

// Exclusive scan of n int32 counts into int32 (or int64) offsets; out[n] = total.
// Implemented in scan.hip.  `flag` (nullable) makes every launch a no-op while
// *flag == 0, so the scan can sit in a step sequence that only sometimes
// rebuilds.  Returns nothing; `tmp` must hold scan_tmp_elems(n) elements.
size_t scan_tmp_elems(int64_t n);
void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* tmp, const int* flag,
                        hipStream_t s, int32_t* max_out = nullptr);
void exclusive_scan_i64(const int32_t* in, int64_t* out, int64_t n, int64_t* tmp, const int* flag,
                        hipStream_t s);

// Fill helpers usable under graph capture (no hipMemsetAsync).
void fill_i32(int32_t* p, int64_t n, int32_t v, const int* flag, hipStream_t s);

}  // namespace gclsph
