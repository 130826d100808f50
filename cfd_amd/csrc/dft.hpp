// dft.hpp -- DFTs of up to 32 points held in registers, for double2 (FP64)
// and float2 (FP32) values.
//
// Sizes 2, 3, 4, 5 and 8 are written out; composite sizes are built at
// compile time as R = P x Q (input n = Q p + q, output k = k1 + P k2):
//   X[k1 + P k2] = sum_q w_Q^{q k2} w_R^{q k1} sum_p w_P^{p k1} x[Q p + q]
// with every index and twiddle a compile-time constant, so a DFT compiles to
// straight-line register code.  Forward w = exp(-2 pi i / R), inverse its
// conjugate; both unnormalised (cft_buff.f:9-13 convention).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "twiddles.hpp"

namespace cfd {

template <typename C>
struct real_of;
template <>
struct real_of<double2> {
  using type = double;
};
template <>
struct real_of<float2> {
  using type = float;
};
template <typename C>
using real_t = typename real_of<C>::type;

template <typename T>
struct cplx_of;
template <>
struct cplx_of<double> {
  using type = double2;
};
template <>
struct cplx_of<float> {
  using type = float2;
};
template <typename T>
using cplx_t = typename cplx_of<T>::type;

template <typename C>
__device__ __forceinline__ C mk(real_t<C> x, real_t<C> y) {
  C c;
  c.x = x;
  c.y = y;
  return c;
}
template <typename C>
__device__ __forceinline__ C vadd(C a, C b) {
  return mk<C>(a.x + b.x, a.y + b.y);
}
template <typename C>
__device__ __forceinline__ C vsub(C a, C b) {
  return mk<C>(a.x - b.x, a.y - b.y);
}
// (-i) b forward, (+i) b inverse
template <bool INV, typename C>
__device__ __forceinline__ C vrot(C b) {
  return INV ? mk<C>(-b.y, b.x) : mk<C>(b.y, -b.x);
}

// compile-time loop: f(std::integral_constant<int, i>) for i in [B, E)
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

template <bool INV, typename C>
__device__ __forceinline__ void vb2(C* v) {
  const C a = v[0], b = v[1];
  v[0] = vadd(a, b);
  v[1] = vsub(a, b);
}

template <bool INV, typename C>
__device__ __forceinline__ void vb3(C* v) {
  using T = real_t<C>;
  const T s = (T)0.86602540378443864676;  // sin(2 pi/3)
  const C t = vadd(v[1], v[2]);
  const C m = mk<C>(v[0].x - (T)0.5 * t.x, v[0].y - (T)0.5 * t.y);
  const C d0 = vsub(v[1], v[2]);
  const C d = vrot<INV>(mk<C>(s * d0.x, s * d0.y));
  v[0] = vadd(v[0], t);
  v[1] = vadd(m, d);
  v[2] = vsub(m, d);
}

template <bool INV, typename C>
__device__ __forceinline__ void vb4(C* v) {
  const C a0 = vadd(v[0], v[2]), a1 = vsub(v[0], v[2]);
  const C b0 = vadd(v[1], v[3]), b1 = vrot<INV>(vsub(v[1], v[3]));
  v[0] = vadd(a0, b0);
  v[2] = vsub(a0, b0);
  v[1] = vadd(a1, b1);
  v[3] = vsub(a1, b1);
}

template <bool INV, typename C>
__device__ __forceinline__ void vb5(C* v) {
  using T = real_t<C>;
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;  // cos(2pi/5), cos(4pi/5)
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;   // sin(2pi/5), sin(4pi/5)
  const C a1 = vadd(v[1], v[4]), b1 = vsub(v[1], v[4]);
  const C a2 = vadd(v[2], v[3]), b2 = vsub(v[2], v[3]);
  const C m1 = mk<C>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
  const C m2 = mk<C>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
  const C n1 = vrot<INV>(mk<C>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  const C n2 = vrot<INV>(mk<C>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = mk<C>(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
  v[1] = vadd(m1, n1);
  v[4] = vsub(m1, n1);
  v[2] = vadd(m2, n2);
  v[3] = vsub(m2, n2);
}

template <bool INV, typename C>
__device__ __forceinline__ void vb8(C* v) {
  // radix-8 as 2 x radix-4 with the internal twiddles w8^k
  using T = real_t<C>;
  const T s = (T)0.70710678118654752440;
  C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  vb4<INV>(e);
  vb4<INV>(o);
  o[1] = INV ? mk<C>(s * (o[1].x - o[1].y), s * (o[1].x + o[1].y)) : mk<C>(s * (o[1].x + o[1].y), s * (o[1].y - o[1].x));
  o[2] = vrot<INV>(o[2]);
  o[3] = INV ? mk<C>(-s * (o[3].x + o[3].y), s * (o[3].x - o[3].y))
             : mk<C>(s * (o[3].y - o[3].x), -s * (o[3].x + o[3].y));
  for (int k = 0; k < 4; ++k) {
    v[k] = vadd(e[k], o[k]);
    v[k + 4] = vsub(e[k], o[k]);
  }
}

// a * w_R^J
template <bool INV, int R, int J, typename C>
__device__ __forceinline__ C twc(C a) {
  using T = real_t<C>;
  constexpr int j = J % R;
  if constexpr (j == 0) {
    return a;
  } else if constexpr (2 * j == R) {
    return mk<C>(-a.x, -a.y);
  } else if constexpr (4 * j == R) {
    return vrot<INV>(a);
  } else if constexpr (4 * j == 3 * R) {
    return vrot<!INV>(a);
  } else {
    constexpr T c = (T)tw<R>::re[j];
    constexpr T s = (T)(INV ? -tw<R>::im[j] : tw<R>::im[j]);
    return mk<C>(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// a * w for a forward twiddle w taken from a table (conjugated for INV)
template <bool INV, typename C>
__device__ __forceinline__ C mul_tw(C a, C w) {
  return INV ? mk<C>(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y) : mk<C>(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

template <int R>
struct dft_split;
template <>
struct dft_split<6> {
  static constexpr int P = 2, Q = 3;
};
template <>
struct dft_split<9> {
  static constexpr int P = 3, Q = 3;
};
template <>
struct dft_split<12> {
  static constexpr int P = 4, Q = 3;
};
template <>
struct dft_split<16> {
  static constexpr int P = 4, Q = 4;
};
template <>
struct dft_split<24> {
  static constexpr int P = 8, Q = 3;
};
template <>
struct dft_split<32> {
  static constexpr int P = 8, Q = 4;
};

// in-place DFT of v[0..R), natural order in and out
template <bool INV, int R, typename C>
__device__ __forceinline__ void dft(C* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    vb2<INV>(v);
  } else if constexpr (R == 3) {
    vb3<INV>(v);
  } else if constexpr (R == 4) {
    vb4<INV>(v);
  } else if constexpr (R == 5) {
    vb5<INV>(v);
  } else if constexpr (R == 8) {
    vb8<INV>(v);
  } else {
    constexpr int P = dft_split<R>::P, Q = dft_split<R>::Q;
    C t[R];
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      C u[P];
      sfor<0, P>([&](auto pc) { u[decltype(pc)::value] = v[Q * decltype(pc)::value + q]; });
      dft<INV, P>(u);
      sfor<0, P>([&](auto kc) {
        constexpr int k1 = decltype(kc)::value;
        t[k1 * Q + q] = twc<INV, R, q * k1>(u[k1]);
      });
    });
    sfor<0, P>([&](auto kc) {
      constexpr int k1 = decltype(kc)::value;
      C u[Q];
      sfor<0, Q>([&](auto qc) { u[decltype(qc)::value] = t[k1 * Q + decltype(qc)::value]; });
      dft<INV, Q>(u);
      sfor<0, Q>([&](auto c) { v[k1 + P * decltype(c)::value] = u[decltype(c)::value]; });
    });
  }
}

}  // namespace cfd
