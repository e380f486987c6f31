// ZDT kernels of the generic path and the step functions.
//
// The chirp-Z transform per frequency bin k (src/zdtfdr.jl:196-321) is
//   Ys[k,j] = post(k,j) w(k) sum_t F[k,t] pre(k,t) chirp(k, j-t),   j < Nr
// with chirp(k,l) = exp(-i pi k l^2 dr/Nf)                       (:140-142)
//      pre(k,t)   = exp( i pi k t (t dr + 2 r0)/Nf)              (:196-198)
//      post(k,j)  = exp( i pi k j^2 dr/Nf) = conj(chirp(k,j))    (:268-270)
// The reference evaluates the convolution over t as a length-Nl circular
// convolution through FFTs (:249-253); the identity chirp(k,l) = chirp(k,-l)
// makes the high lags of V (:182-184) the negative lags, so the result equals
// the linear convolution above for any Nl >= Nt+Nr-1.  czt_reinsch (Nt <= 64)
// and czt_reinsch_chunks (Nt > 64) evaluate it directly (about Nt operations
// per output), which costs less than the two length-Nl FFTs, needs no V, and
// is exact up to float rounding.  The step functions (preprocess!,
// convolve!, postprocess!) keep the reference's FFT data flow on the
// materialized Nk x Nl matrix Y for API parity.
#include <hip/hip_runtime.h>

#include "zdt_kernels.h"

namespace zdt {

// exp(2 pi i rev) in double for any rev (reduced first)
__device__ __forceinline__ c64 cis64(double rev) {
  const double f = rev - __builtin_rint(rev);
  double s, c;
  sincospi(2.0 * f, &s, &c);
  return {c, s};
}

// exp(2 pi i (n1*h1 + n2*h2)) in float, each product reduced mod 1 in double
__device__ __forceinline__ c32 cis_sum2(double n1, double h1, double n2, double h2) {
  double p1 = n1 * h1;
  p1 -= __builtin_rint(p1);
  double p2 = n2 * h2;
  p2 -= __builtin_rint(p2);
  double p = p1 + p2;
  p -= __builtin_rint(p);
  return cis_rev(static_cast<float>(p));
}

__device__ __forceinline__ float window_value(int wkind, const float* wcustom, int64_t k, int64_t Nf) {
  if (wkind == ZDT_WINDOW_HAMMING) {
    // 0.53836 + 0.46164 cospi(2k/Nf)  (src/zdtfdr.jl:303-305)
    double p = static_cast<double>(k) / static_cast<double>(Nf);
    p -= __builtin_rint(p);
    const c32 c = cis_rev(static_cast<float>(p));
    return 0.53836f + 0.46164f * c.x;
  }
  if (wkind == ZDT_WINDOW_CUSTOM) return wcustom[k];
  return 1.0f;
}

// ---------------------------------------------------------------------------
// Vandermonde form for Nt <= 64: by Bluestein's identity the sum above is
//   Ys[k,j] = w(k) sum_t a_t z_j^t,  a_t = F[k,t] e^{2 pi i k t r0/Nf},
//   z_j = gamma^j, gamma = e^{2 pi i k dr/Nf},
// a polynomial in z_j.  A thread keeps its bin's coefficients in registers
// and writes every Ys[k,j] once.  The rates are taken in mirrored pairs: with the
// centre rate rc = r0 + (Nr-1) dr/2 folded into the coefficients
// (c_t = F[k,t] e^{2 pi i k t rc/Nf}), rates j and Nr-1-j are the sums at
// z = e^{i theta} and conj(z), theta = 2 pi k (r0 + j dr - rc)/Nf, and one
// Reinsch-modified Goertzel recurrence over the real and the imaginary parts
// gives both (C. Reinsch, 1967; J. Oliver, "An error analysis of the modified
// Clenshaw method", 1977).  With s = sign(cos theta), lambda = 2 (cos theta - s),
//   D_t = c_t + lambda U_{t+1} + s D_{t+1},  U_t = D_t + s U_{t+1}  (t = Nt-1 .. 1),
//   sum c_t cos(t theta) = c_0 + (lambda/2) U_1 + s D_1,  sum c_t sin(t theta) = U_1 sin theta,
// (evaluated with s folded into the state, so that every update is an
// in-place add or FMA), six operations per term for the two rates against
// eight for two complex Horner
// evaluations.  z comes from an FP64 recurrence started at an exactly
// evaluated value, and lambda from the FP64 cos theta, so no accuracy is lost
// near theta = 0 or pi.  The middle rate of an odd Nr is its own partner.
template <int NT>
__global__ __launch_bounds__(256) void czt_reinsch(CztArgs a) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= a.Nk) return;
  const double kd = static_cast<double>(k);
  const float w = window_value(a.wkind, a.wcustom, k, a.Nf) * a.scale;
  float cr[NT], ci[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < a.Nt) {
      c32 f = a.F[k + static_cast<int64_t>(t) * a.ldF];
      if (a.hc != 0.0) f = cmul(f, cis_nh(kd * static_cast<double>(t), a.hc));  // uniform
      cr[t] = f.x * w;
      ci[t] = f.y * w;
    } else {
      cr[t] = 0.0f;
      ci[t] = 0.0f;
    }
  }
  const c64 g = cis64(2.0 * kd * a.hd);  // gamma = e^{2 pi i k dr/Nf}
  c64 z = cis64(kd * a.hz);              // e^{2 pi i k (r0 - rc)/Nf}
  c32* dst = a.Ys + k;
  const int npair = (a.Nr + 1) / 2;  // j < npair; partner Nr-1-j
  for (int j = 0; j < npair; ++j) {
    const int jp = a.Nr - 1 - j;
    // the recurrence with s folded into the state (D~_t = s^t D_t,
    // U~_t = s^t U_t): every update accumulates in place
    const float s = z.x >= 0.0 ? 1.0f : -1.0f;
    const float lamp = static_cast<float>(2.0 * (z.x - static_cast<double>(s))) * s;
    const float ssn = static_cast<float>(z.y) * s;
    constexpr bool odd = (NT - 1) & 1;
    float Dr = odd ? s * cr[NT - 1] : cr[NT - 1], Di = odd ? s * ci[NT - 1] : ci[NT - 1];
    float Ur = Dr, Ui = Di;
#pragma unroll
    for (int t = NT - 2; t >= 1; --t) {
      if (t & 1) {
        Dr = fmaf(s, cr[t], Dr);
        Di = fmaf(s, ci[t], Di);
      } else {
        Dr += cr[t];
        Di += ci[t];
      }
      Dr = fmaf(lamp, Ur, Dr);
      Di = fmaf(lamp, Ui, Di);
      Ur += Dr;
      Ui += Di;
    }
    const float hl = 0.5f * lamp;
    const float Cr = fmaf(hl, Ur, cr[0]) + Dr;
    const float Ci = fmaf(hl, Ui, ci[0]) + Di;
    const float Sr = ssn * Ur, Si = ssn * Ui;
    dst[static_cast<int64_t>(j) * a.ldY] = {Cr - Si, Sr + Ci};
    if (jp != j) dst[static_cast<int64_t>(jp) * a.ldY] = {Cr + Si, Ci - Sr};
    z = cmul(z, g);
  }
}

// Nt > 64: the sum in chunks of 64 time steps, sum_t c_t z^t =
// sum_q z^{64q} P_q(z) with P_q the polynomial of chunk q, evaluated for both
// rates of a pair by the recurrence of czt_reinsch; z^{64q} (conj for the
// partner) comes from an FP64 recurrence over the rates.  Chunk 0 writes Ys,
// later chunks add to it (one read-modify-write of Ys per further chunk).
__global__ __launch_bounds__(256) void czt_reinsch_chunks(CztArgs a) {
  constexpr int NT = 64;
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= a.Nk) return;
  const double kd = static_cast<double>(k);
  const float w = window_value(a.wkind, a.wcustom, k, a.Nf) * a.scale;
  const c64 g = cis64(2.0 * kd * a.hd);  // gamma = e^{2 pi i k dr/Nf}
  c32* dst = a.Ys + k;
  const int npair = (a.Nr + 1) / 2;  // j < npair; partner Nr-1-j
  for (int t0 = 0; t0 < a.Nt; t0 += NT) {
    float cr[NT], ci[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      const int t = t0 + u;
      if (t < a.Nt) {
        c32 f = a.F[k + static_cast<int64_t>(t) * a.ldF];
        if (a.hc != 0.0) f = cmul(f, cis_nh(kd * static_cast<double>(t), a.hc));  // uniform
        cr[u] = f.x * w;
        ci[u] = f.y * w;
      } else {
        cr[u] = 0.0f;
        ci[u] = 0.0f;
      }
    }
    const double td = static_cast<double>(t0);
    c64 z = cis64(kd * a.hz);               // e^{2 pi i k (r0 - rc)/Nf}
    c64 zq = cis64(kd * a.hz * td);         // z^{t0} of rate 0
    const c64 gq = cis64(2.0 * kd * a.hd * td);  // gamma^{t0}
    for (int j = 0; j < npair; ++j) {
      const int jp = a.Nr - 1 - j;
      const float s = z.x >= 0.0 ? 1.0f : -1.0f;
      const float lamp = static_cast<float>(2.0 * (z.x - static_cast<double>(s))) * s;
      const float ssn = static_cast<float>(z.y) * s;
      float Dr = s * cr[NT - 1], Di = s * ci[NT - 1];  // NT - 1 is odd
      float Ur = Dr, Ui = Di;
#pragma unroll
      for (int t = NT - 2; t >= 1; --t) {
        if (t & 1) {
          Dr = fmaf(s, cr[t], Dr);
          Di = fmaf(s, ci[t], Di);
        } else {
          Dr += cr[t];
          Di += ci[t];
        }
        Dr = fmaf(lamp, Ur, Dr);
        Di = fmaf(lamp, Ui, Di);
        Ur += Dr;
        Ui += Di;
      }
      const float hl = 0.5f * lamp;
      const float Cr = fmaf(hl, Ur, cr[0]) + Dr;
      const float Ci = fmaf(hl, Ui, ci[0]) + Di;
      const float Sr = ssn * Ur, Si = ssn * Ui;
      const c32 q = {static_cast<float>(zq.x), static_cast<float>(zq.y)};
      c32 yz = cmul(c32{Cr - Si, Sr + Ci}, q);
      c32 yc = cmulc(c32{Cr + Si, Ci - Sr}, q);
      c32* dj = dst + static_cast<int64_t>(j) * a.ldY;
      c32* dp = dst + static_cast<int64_t>(jp) * a.ldY;
      if (t0 > 0) {
        yz = cadd(yz, *dj);
        if (jp != j) yc = cadd(yc, *dp);
      }
      *dj = yz;
      if (jp != j) *dp = yc;
      z = cmul(z, g);
      zq = cmul(zq, gq);
    }
  }
}

int launch_czt(const CztArgs& a, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((a.Nk + 255) / 256);
  if (a.Nt > 64) {
    ProfScope ps(st, "czt_reinsch_chunks",
                 8.0 * static_cast<double>(a.Nk) * (a.Nt + a.Nr * (2.0 * ((a.Nt + 63) / 64) - 1.0)));
    czt_reinsch_chunks<<<grid, 256, 0, st>>>(a);
    ZDT_CHECK_LAUNCH("czt_reinsch_chunks");
    return ZDT_OK;
  }
  ProfScope ps(st, "czt_reinsch", 8.0 * static_cast<double>(a.Nk) * (a.Nt + a.Nr));
  if (a.Nt <= 2) {
    czt_reinsch<2><<<grid, 256, 0, st>>>(a);
  } else if (a.Nt <= 4) {
    czt_reinsch<4><<<grid, 256, 0, st>>>(a);
  } else if (a.Nt <= 8) {
    czt_reinsch<8><<<grid, 256, 0, st>>>(a);
  } else if (a.Nt <= 16) {
    czt_reinsch<16><<<grid, 256, 0, st>>>(a);
  } else if (a.Nt <= 32) {
    czt_reinsch<32><<<grid, 256, 0, st>>>(a);
  } else {
    czt_reinsch<64><<<grid, 256, 0, st>>>(a);
  }
  ZDT_CHECK_LAUNCH("czt_reinsch");
  return ZDT_OK;
}

// ---------------------------------------------------------------------------
// fftfdr (src/fftfdr.jl:86-96,126-148) for an arbitrary list of rates:
//   Ys[k,j] = sum_t F[k,t] e^{2 pi i k t rates[j]/Nf}.
// Per bin and rate z = e^{2 pi i k rates[j]/Nf} and z^16 are evaluated
// directly (k*rate/Nf reduced in FP64), and the polynomial is evaluated as
// 16-term Horner blocks combined by Horner in z^16, so no power of a rounded
// z above the 15th enters the sum.  For Nt > 64 every term's phase is
// evaluated directly.
template <int NT>
__global__ __launch_bounds__(256) void czt_rates(const c32* __restrict__ F, int64_t ldF, c32* __restrict__ Ys,
                                                 int64_t ldY, int64_t Nk, int64_t Nf, int Nt,
                                                 const double* __restrict__ rh, int Nr) {
  constexpr int B = NT < 16 ? NT : 16;
  constexpr int NB = NT / B;
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= Nk) return;
  const double kd = static_cast<double>(k);
  c32 c[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) c[t] = (t < Nt) ? F[k + static_cast<int64_t>(t) * ldF] : c32{0.0f, 0.0f};
  for (int j = 0; j < Nr; ++j) {
    const c32 zf = cis_nh(kd, rh[j]);                          // rh[j] = rates[j]/Nf
    const c32 zBf = cis_nh(kd * static_cast<double>(B), rh[j]);
    c32 acc = {0.0f, 0.0f};
#pragma unroll
    for (int b = NB - 1; b >= 0; --b) {
      c32 p = c[b * B + B - 1];
#pragma unroll
      for (int u = B - 2; u >= 0; --u) p = cfma(c[b * B + u], p, zf);
      acc = (b == NB - 1) ? p : cfma(p, acc, zBf);
    }
    Ys[k + static_cast<int64_t>(j) * ldY] = acc;
  }
}

__global__ __launch_bounds__(256) void czt_rates_any(const c32* __restrict__ F, int64_t ldF, c32* __restrict__ Ys,
                                                     int64_t ldY, int64_t Nk, int Nt, const double* __restrict__ rh,
                                                     int Nr) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= Nk) return;
  const double kd = static_cast<double>(k);
  for (int j = 0; j < Nr; ++j) {
    c32 acc = {0.0f, 0.0f};
    for (int t = 0; t < Nt; ++t)
      acc = cfma(acc, F[k + static_cast<int64_t>(t) * ldF], cis_nh(kd * static_cast<double>(t), rh[j]));
    Ys[k + static_cast<int64_t>(j) * ldY] = acc;
  }
}

int launch_czt_rates(const c32* F, int64_t ldF, c32* Ys, int64_t ldY, int64_t Nk, int64_t Nf, int Nt,
                     const double* rates_over_nf, int Nr, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((Nk + 255) / 256);
  ProfScope ps(st, "czt_rates", 8.0 * static_cast<double>(Nk) * (Nt + Nr));
  const double* rh = rates_over_nf;
  if (Nt <= 4) {
    czt_rates<4><<<grid, 256, 0, st>>>(F, ldF, Ys, ldY, Nk, Nf, Nt, rh, Nr);
  } else if (Nt <= 16) {
    czt_rates<16><<<grid, 256, 0, st>>>(F, ldF, Ys, ldY, Nk, Nf, Nt, rh, Nr);
  } else if (Nt <= 32) {
    czt_rates<32><<<grid, 256, 0, st>>>(F, ldF, Ys, ldY, Nk, Nf, Nt, rh, Nr);
  } else if (Nt <= 64) {
    czt_rates<64><<<grid, 256, 0, st>>>(F, ldF, Ys, ldY, Nk, Nf, Nt, rh, Nr);
  } else {
    czt_rates_any<<<grid, 256, 0, st>>>(F, ldF, Ys, ldY, Nk, Nt, rh, Nr);
  }
  ZDT_CHECK_LAUNCH("czt_rates");
  return ZDT_OK;
}

// fdshift (src/fftfdr.jl:104-111): Y[k,t] = F[k,t] e^{2 pi i k t rate/Nf}
__global__ __launch_bounds__(256) void shift_phase(const c32* __restrict__ F, int64_t ldF, c32* __restrict__ Y,
                                                   int64_t ldY, int64_t Nk, int Nt, double rh) {
  const int64_t total = Nk * Nt;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = g % Nk, t = g / Nk;
    Y[k + t * ldY] = cmul(F[k + t * ldF], cis_nh(static_cast<double>(k) * static_cast<double>(t), rh));
  }
}

int launch_shift_phase(const c32* F, int64_t ldF, c32* Y, int64_t ldY, int64_t Nk, int Nt, double rate_over_nf,
                       hipStream_t st) {
  int64_t g = (Nk * Nt + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  shift_phase<<<static_cast<unsigned>(g), 256, 0, st>>>(F, ldF, Y, ldY, Nk, Nt, rate_over_nf);
  ZDT_CHECK_LAUNCH("shift_phase");
  return ZDT_OK;
}


// ---------------------------------------------------------------------------
// step functions (materialized Y, src/zdtfdr.jl:221-321)
// ---------------------------------------------------------------------------

// preprocess!: Y[:, :Nt] = F .* prephase, Y[:, Nt:] = 0   (src/zdtfdr.jl:221-235)
__global__ __launch_bounds__(256) void step_prephase(const c32* __restrict__ F, int64_t ldF,
                                                     c32* __restrict__ Y, int64_t ldY, int64_t Nk,
                                                     int Nt, int Nl, double hd, double h0) {
  const int64_t total = Nk * Nl;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = g % Nk;
    const int t = static_cast<int>(g / Nk);
    c32 v = {0.0f, 0.0f};
    if (t < Nt) {
      const double kd = static_cast<double>(k), td = static_cast<double>(t);
      v = cmul(F[k + t * ldF], cis_sum2(kd * td * td, hd, kd * td, h0));
    }
    Y[k + t * ldY] = v;
  }
}

// chirp kernel before its FFT, in FP64 (src/zdtfdr.jl:173-185):
// V0[k,l] = chirp(k,l) for l < Nr, chirp(k, Nl-l) for l >= Nl-Nt+1, else 0
__global__ __launch_bounds__(256) void step_vgen(c64* __restrict__ V, int64_t Nk, int Nt, int Nr, int Nl,
                                                 double hd) {
  const int64_t total = Nk * Nl;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = g % Nk;
    const int l = static_cast<int>(g / Nk);
    c64 v = {0.0, 0.0};
    int lag = -1;
    if (l < Nr) {
      lag = l;
    } else if (l >= Nl - Nt + 1) {
      lag = Nl - l;
    }
    if (lag >= 0) {
      const double kd = static_cast<double>(k), ld = static_cast<double>(lag);
      v = cis64(-(kd * ld * ld) * hd);
    }
    V[g] = v;
  }
}

__global__ __launch_bounds__(256) void c64_to_c32(const c64* __restrict__ in, c32* __restrict__ out,
                                                  int64_t n) {
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < n;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    out[g] = {static_cast<float>(in[g].x), static_cast<float>(in[g].y)};
  }
}

// Y .*= V  (src/zdtfdr.jl:251)
__global__ __launch_bounds__(256) void step_vmul(c32* __restrict__ Y, const c32* __restrict__ V, int64_t n) {
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < n;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    Y[g] = cmul(Y[g], V[g]);
  }
}

// postprocess!: Ys .*= postphase .* w  (src/zdtfdr.jl:297-321)
__global__ __launch_bounds__(256) void step_postphase(c32* __restrict__ Y, int64_t ldY, int64_t Nk,
                                                      int64_t Nf, int Nr, double hd, int wkind,
                                                      const float* __restrict__ wcustom, float scale) {
  const int64_t total = Nk * Nr;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = g % Nk;
    const int j = static_cast<int>(g / Nk);
    const double kd = static_cast<double>(k), jd = static_cast<double>(j);
    double p = (kd * jd * jd) * hd;
    p -= __builtin_rint(p);
    const c32 post = cis_rev(static_cast<float>(p));
    const float w = window_value(wkind, wcustom, k, Nf) * scale;
    c32 v = cmul(Y[k + j * ldY], post);
    Y[k + j * ldY] = {v.x * w, v.y * w};
  }
}

static unsigned grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

int launch_step_prephase(const c32* F, int64_t ldF, c32* Y, int64_t ldY, int64_t Nk, int Nt, int Nl,
                         double hd, double h0, hipStream_t st) {
  ProfScope ps(st, "step_prephase", 8.0 * static_cast<double>(Nk) * (Nt + Nl));
  step_prephase<<<grid_for(Nk * Nl), 256, 0, st>>>(F, ldF, Y, ldY, Nk, Nt, Nl, hd, h0);
  ZDT_CHECK_LAUNCH("step_prephase");
  return ZDT_OK;
}

int launch_step_vgen(c64* V, int64_t Nk, int Nt, int Nr, int Nl, double hd, hipStream_t st) {
  step_vgen<<<grid_for(Nk * Nl), 256, 0, st>>>(V, Nk, Nt, Nr, Nl, hd);
  ZDT_CHECK_LAUNCH("step_vgen");
  return ZDT_OK;
}

int launch_c64_to_c32(const c64* in, c32* out, int64_t n, hipStream_t st) {
  c64_to_c32<<<grid_for(n), 256, 0, st>>>(in, out, n);
  ZDT_CHECK_LAUNCH("c64_to_c32");
  return ZDT_OK;
}

int launch_step_vmul(c32* Y, const c32* V, int64_t n, hipStream_t st) {
  ProfScope ps(st, "step_vmul", 24.0 * static_cast<double>(n));
  step_vmul<<<grid_for(n), 256, 0, st>>>(Y, V, n);
  ZDT_CHECK_LAUNCH("step_vmul");
  return ZDT_OK;
}

int launch_step_postphase(c32* Y, int64_t ldY, int64_t Nk, int64_t Nf, int Nr, double hd, int wkind,
                          const float* wcustom, float scale, hipStream_t st) {
  ProfScope ps(st, "step_postphase", 16.0 * static_cast<double>(Nk * Nr));
  step_postphase<<<grid_for(Nk * Nr), 256, 0, st>>>(Y, ldY, Nk, Nf, Nr, hd, wkind, wcustom, scale);
  ZDT_CHECK_LAUNCH("step_postphase");
  return ZDT_OK;
}

}  // namespace zdt
