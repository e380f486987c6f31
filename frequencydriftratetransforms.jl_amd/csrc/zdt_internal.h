// Internal declarations shared by the ZDT HIP sources (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/zdt.h"

namespace zdt {

// ---------------------------------------------------------------------------
// error reporting (thread-local message, see zdt_last_error)
// ---------------------------------------------------------------------------
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define ZDT_TRY_HIP(call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return ::zdt::hip_fail(e_, #call); \
  } while (0)

#define ZDT_TRY(call)           \
  do {                          \
    int rc_ = (call);           \
    if (rc_ != ZDT_OK) return rc_; \
  } while (0)

// launch-error check after a <<<>>> launch
#define ZDT_CHECK_LAUNCH(what) ZDT_TRY_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// complex helpers (host+device)
// ---------------------------------------------------------------------------
template <typename T>
struct cplx {
  T x, y;
};
using c32 = cplx<float>;
using c64 = cplx<double>;

template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // a*conj(b)
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
  return {a.x, -a.y};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
  return {a.x * s, a.y * s};
}
// a + b*c
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> c) {
  return {a.x + b.x * c.x - b.y * c.y, a.y + b.x * c.y + b.y * c.x};
}

// ---------------------------------------------------------------------------
// exact phases
//
// Every phase of the transform has the form exp(i*pi*n*q) with n an integer
// and q = (dyadic Float32 rate quantity)/Nf.  The host hands the kernels
// q/2 in double (exact when Nf is a power of two, relative error 2^-53
// otherwise).  The device forms n*(q/2) in double (n < 2^53 is exact), takes
// the fractional part, and evaluates cos/sin(2*pi*f) for f in [-1/2, 1/2] in
// float with a minimax polynomial.  The result is accurate to about one
// float ulp regardless of |n*q|, unlike rounding n*q to Float32 first.
// ---------------------------------------------------------------------------

// cos/sin(2*pi*u) for |u| <= 1/8: sin = u*P(u^2) (degree 3), cos = Q(u^2)
// (degree 4), minimax coefficients from a Remez exchange (gen_sincos_coeffs.py
// beside this file); float32 evaluation error < 9e-8 absolute.
__device__ __forceinline__ void sincos_2pi_small(float u, float* s, float* c) {
  const float z = u * u;
  float ps = -7.54020691e+01f;
  ps = __builtin_fmaf(ps, z, 8.15925522e+01f);
  ps = __builtin_fmaf(ps, z, -4.13416634e+01f);
  ps = __builtin_fmaf(ps, z, 6.28318548e+00f);
  float pc = 5.92204056e+01f;
  pc = __builtin_fmaf(pc, z, -8.54428558e+01f);
  pc = __builtin_fmaf(pc, z, 6.49393158e+01f);
  pc = __builtin_fmaf(pc, z, -1.97392082e+01f);
  pc = __builtin_fmaf(pc, z, 1.0f);
  *s = u * ps;
  *c = pc;
}

// exp(2*pi*i*f) for any float f (reduced to the nearest quarter first).
__device__ __forceinline__ c32 cis_rev(float f) {
  const float q = __builtin_rintf(4.0f * f);
  const float u = __builtin_fmaf(q, -0.25f, f);  // exact
  float s, c;
  sincos_2pi_small(u, &s, &c);
  const int iq = static_cast<int>(q) & 3;
  c32 r;
  // rotate by iq quarter turns
  if (iq == 0) {
    r = {c, s};
  } else if (iq == 1) {
    r = {-s, c};
  } else if (iq == 2) {
    r = {-c, -s};
  } else {
    r = {s, -c};
  }
  return r;
}

// exp(2*pi*i*(n*h)) with n an integer-valued double and h a double; the
// product is reduced modulo 1 in double before the float evaluation.
__device__ __forceinline__ c32 cis_nh(double n, double h) {
  const double p = n * h;
  const double f = p - __builtin_rint(p);
  return cis_rev(static_cast<float>(f));
}

// ---------------------------------------------------------------------------
// kernel profiler: when a workspace carries ZDT_FLAG_PROFILE, the entry points
// install its profiler for the calling thread and every launch site records a
// HIP event pair on its stream, tagged with the kernel's name and the
// algorithmic HBM bytes of that launch.
// ---------------------------------------------------------------------------
struct Profiler;
extern thread_local Profiler* g_prof;
void prof_begin(hipStream_t st, const char* name, double bytes);
void prof_end(hipStream_t st);

struct ProfScope {
  hipStream_t st;
  bool on;
  ProfScope(hipStream_t s, const char* name, double bytes) : st(s), on(g_prof != nullptr) {
    if (on) prof_begin(st, name, bytes);
  }
  ~ProfScope() {
    if (on) prof_end(st);
  }
};

// the profiler of a workspace (null without ZDT_FLAG_PROFILE), and a scope
// that installs a profiler for the calling thread (launch sites outside the
// workspace entry points: the long-integration search's hit extraction)
Profiler* ws_profiler(const zdt_ws* ws);
struct ProfUse {
  Profiler* prev;
  explicit ProfUse(Profiler* p) : prev(g_prof) { g_prof = p; }
  ~ProfUse() { g_prof = prev; }
};

}  // namespace zdt
