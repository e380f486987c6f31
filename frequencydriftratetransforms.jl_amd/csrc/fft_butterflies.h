// Small DFT butterflies (forward, W_R = exp(-2 pi i/R)) shared by the FFT
// kernels.  Inverse transforms use conj(DFT(conj(x))).
#pragma once

#include "zdt_internal.h"

namespace zdt {

template <typename T>
__device__ __forceinline__ cplx<T> mul_neg_i(cplx<T> a) {  // a * (-i)
  return {a.y, -a.x};
}

template <typename T>
__device__ __forceinline__ void dft2(cplx<T>& a, cplx<T>& b) {
  cplx<T> t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>* v) {
  cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  cplx<T> t2 = cadd(v[1], v[3]), t3 = mul_neg_i(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

// length-2H DFT from two length-H DFTs (even/odd split) with W_{2H}^k
template <typename T, int H>
__device__ __forceinline__ void combine_even_odd(const cplx<T>* e, const cplx<T>* o, cplx<T>* out,
                                                 const double* cw, const double* sw) {
#pragma unroll
  for (int k = 0; k < H; ++k) {
    cplx<T> ow;
    if (k == 0) {
      ow = o[0];
    } else if (2 * k == H) {
      ow = mul_neg_i(o[k]);
    } else {
      const cplx<T> w = {static_cast<T>(cw[k]), static_cast<T>(-sw[k])};
      ow = cmul(o[k], w);
    }
    out[k] = cadd(e[k], ow);
    out[k + H] = csub(e[k], ow);
  }
}

template <typename T>
__device__ __forceinline__ void dft8(cplx<T>* v) {
  constexpr double c8[4] = {1, 0.70710678118654757, 0, -0.70710678118654757};
  constexpr double s8[4] = {0, 0.70710678118654757, 1, 0.70710678118654757};
  cplx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e);
  dft4(o);
  combine_even_odd<T, 4>(e, o, v, c8, s8);
}

template <typename T>
__device__ __forceinline__ void dft16(cplx<T>* v) {
  constexpr double c16[8] = {1, 0.92387953251128674, 0.70710678118654757, 0.38268343236508978,
                             0, -0.38268343236508978, -0.70710678118654757, -0.92387953251128674};
  constexpr double s16[8] = {0, 0.38268343236508978, 0.70710678118654757, 0.92387953251128674,
                             1, 0.92387953251128674, 0.70710678118654757, 0.38268343236508978};
  cplx<T> e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft8(e);
  dft8(o);
  combine_even_odd<T, 8>(e, o, v, c16, s16);
}

template <typename T>
__device__ __forceinline__ void dft32(cplx<T>* v) {
  constexpr double c32[16] = {1, 0.98078528040323043, 0.92387953251128674, 0.83146961230254524, 0.70710678118654757, 0.55557023301960229, 0.38268343236508984, 0.19509032201612833, 6.123233995736766e-17, -0.19509032201612819, -0.38268343236508973, -0.55557023301960196, -0.70710678118654746, -0.83146961230254535, -0.92387953251128674, -0.98078528040323043};
  constexpr double s32[16] = {0, 0.19509032201612825, 0.38268343236508978, 0.55557023301960218, 0.70710678118654746, 0.83146961230254524, 0.92387953251128674, 0.98078528040323043, 1, 0.98078528040323043, 0.92387953251128674, 0.83146961230254546, 0.70710678118654757, 0.55557023301960218, 0.38268343236508989, 0.19509032201612861};
  cplx<T> e[16], o[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft16(e);
  dft16(o);
  combine_even_odd<T, 16>(e, o, v, c32, s32);
}

template <int R>
struct OddTab;
template <>
struct OddTab<3> {
  static constexpr double c[3] = {1, -0.5, -0.5};
  static constexpr double s[3] = {0, 0.86602540378443865, -0.86602540378443865};
};
template <>
struct OddTab<5> {
  static constexpr double c[5] = {1, 0.30901699437494745, -0.80901699437494734, -0.80901699437494756,
                                  0.30901699437494723};
  static constexpr double s[5] = {0, 0.95105651629515353, 0.58778525229247325, -0.58778525229247303,
                                  -0.95105651629515364};
};
template <>
struct OddTab<7> {
  static constexpr double c[7] = {1, 0.62348980185873359, -0.22252093395631434, -0.90096886790241903,
                                  -0.90096886790241915, -0.22252093395631459, 0.62348980185873337};
  static constexpr double s[7] = {0, 0.7818314824680298, 0.97492791218182362, 0.43388373911755823,
                                  -0.43388373911755801, -0.97492791218182362, -0.78183148246802991};
};
template <>
struct OddTab<11> {
  static constexpr double c[11] = {1, 0.84125353283118121, 0.41541501300188644, -0.142314838273285,
                                   -0.65486073394528499, -0.95949297361449737, -0.95949297361449748,
                                   -0.65486073394528521, -0.14231483827328523, 0.41541501300188605,
                                   0.84125353283118121};
  static constexpr double s[11] = {0, 0.54064081745559756, 0.90963199535451833, 0.9898214418809328,
                                   0.75574957435425827, 0.28173255684142967, -0.28173255684142939,
                                   -0.75574957435425816, -0.98982144188093268, -0.90963199535451855,
                                   -0.54064081745559744};
};
template <>
struct OddTab<13> {
  static constexpr double c[13] = {1, 0.88545602565320991, 0.56806474673115592, 0.12053668025532301,
                                   -0.35460488704253545, -0.74851074817110119, -0.97094181742605201,
                                   -0.97094181742605212, -0.7485107481711013, -0.3546048870425359,
                                   0.1205366802553232, 0.56806474673115481, 0.88545602565321002};
  static constexpr double s[13] = {0, 0.46472317204376851, 0.82298386589365635, 0.99270887409805397,
                                   0.93501624268541483, 0.66312265824079519, 0.23931566428755768,
                                   -0.23931566428755743, -0.66312265824079497, -0.93501624268541472,
                                   -0.99270887409805397, -0.82298386589365702, -0.4647231720437684};
};

// odd-length DFT in the symmetric form: for k in 1..(R-1)/2
//   a_k = x0 + sum_m cos(2 pi m k/R) (x_m + x_{R-m}),  b_k = sum_m sin(...) (x_m - x_{R-m})
//   y_k = a_k - i b_k,  y_{R-k} = a_k + i b_k
template <typename T, int R>
__device__ __forceinline__ void dft_odd(cplx<T>* v) {
  constexpr int H = (R - 1) / 2;
  cplx<T> p[H], m[H];
  cplx<T> y0 = v[0];
#pragma unroll
  for (int i = 1; i <= H; ++i) {
    p[i - 1] = cadd(v[i], v[R - i]);
    m[i - 1] = csub(v[i], v[R - i]);
    y0 = cadd(y0, p[i - 1]);
  }
  cplx<T> out[R];
  out[0] = y0;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    cplx<T> a = v[0], b = {0, 0};
#pragma unroll
    for (int i = 1; i <= H; ++i) {
      const int idx = (i * k) % R;
      const T c = static_cast<T>(OddTab<R>::c[idx]);
      const T s = static_cast<T>(OddTab<R>::s[idx]);
      a = {a.x + c * p[i - 1].x, a.y + c * p[i - 1].y};
      b = {b.x + s * m[i - 1].x, b.y + s * m[i - 1].y};
    }
    // -i*b = (b.y, -b.x)
    out[k] = {a.x + b.y, a.y - b.x};
    out[R - k] = {a.x - b.y, a.y + b.x};
  }
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = out[i];
}

template <typename T, int R>
__device__ __forceinline__ void dft(cplx<T>* v) {
  if constexpr (R == 2) {
    dft2(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4(v);
  } else if constexpr (R == 8) {
    dft8(v);
  } else if constexpr (R == 16) {
    dft16(v);
  } else if constexpr (R == 32) {
    dft32(v);
  } else {
    dft_odd<T, R>(v);
  }
}

// ---------------------------------------------------------------- direction-templated
// INV = false: W_R = exp(-2 pi i/R); INV = true: W_R = exp(+2 pi i/R) (unnormalized inverse)

template <bool INV, typename T>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) {  // a * (-i) forward, a * (+i) inverse
  return INV ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

template <bool INV, typename T>
__device__ __forceinline__ void dft4d(cplx<T>* v) {
  cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  cplx<T> t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <bool INV, typename T, int H>
__device__ __forceinline__ void combine_d(const cplx<T>* e, const cplx<T>* o, cplx<T>* out, const double* cw,
                                          const double* sw) {
#pragma unroll
  for (int k = 0; k < H; ++k) {
    cplx<T> ow;
    if (k == 0) {
      ow = o[0];
    } else if (2 * k == H) {
      ow = mul_mi<INV>(o[k]);
    } else {
      const cplx<T> w = {static_cast<T>(cw[k]), static_cast<T>(INV ? sw[k] : -sw[k])};
      ow = cmul(o[k], w);
    }
    out[k] = cadd(e[k], ow);
    out[k + H] = csub(e[k], ow);
  }
}

template <bool INV, typename T>
__device__ __forceinline__ void dft8d(cplx<T>* v) {
  constexpr double c8[4] = {1, 0.70710678118654757, 0, -0.70710678118654757};
  constexpr double s8[4] = {0, 0.70710678118654757, 1, 0.70710678118654757};
  cplx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  dft4d<INV>(e);
  dft4d<INV>(o);
  combine_d<INV, T, 4>(e, o, v, c8, s8);
}

template <bool INV, typename T>
__device__ __forceinline__ void dft16d(cplx<T>* v) {
  constexpr double c16[8] = {1, 0.92387953251128674, 0.70710678118654757, 0.38268343236508978,
                             0, -0.38268343236508978, -0.70710678118654757, -0.92387953251128674};
  constexpr double s16[8] = {0, 0.38268343236508978, 0.70710678118654757, 0.92387953251128674,
                             1, 0.92387953251128674, 0.70710678118654757, 0.38268343236508978};
  cplx<T> e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft8d<INV>(e);
  dft8d<INV>(o);
  combine_d<INV, T, 8>(e, o, v, c16, s16);
}

template <bool INV, typename T, int R>
__device__ __forceinline__ void dftd(cplx<T>* v) {
  if constexpr (R == 2) {
    cplx<T> t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
  } else if constexpr (R == 4) {
    dft4d<INV>(v);
  } else if constexpr (R == 8) {
    dft8d<INV>(v);
  } else {
    static_assert(R == 16, "radix");
    dft16d<INV>(v);
  }
}

}  // namespace zdt
