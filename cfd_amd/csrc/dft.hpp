// dft.hpp -- DFTs of up to 32 points held in registers.
//
// Sizes 2, 3, 4, 5 and 8 are the butterflies of fft.hpp; composite sizes are
// built at compile time as R = P x Q (input n = Q p + q, output k = k1 + P k2):
//   X[k1 + P k2] = sum_q w_Q^{q k2} w_R^{q k1} sum_p w_P^{p k1} x[Q p + q]
// with every index and twiddle a compile-time constant, so a DFT compiles to
// straight-line register code.  Forward w = exp(-2 pi i / R), inverse its
// conjugate; both unnormalised (cft_buff.f:9-13 convention).
#pragma once
#include <type_traits>

#include "fft.hpp"
#include "twiddles.hpp"

namespace cfd {

// compile-time loop: f(std::integral_constant<int, i>) for i in [B, E)
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// a * w_R^J
template <bool INV, int R, int J>
__device__ __forceinline__ double2 twc(double2 a) {
  constexpr int j = J % R;
  if constexpr (j == 0) {
    return a;
  } else if constexpr (2 * j == R) {
    return make_double2(-a.x, -a.y);
  } else if constexpr (4 * j == R) {
    return rot<INV>(a);
  } else if constexpr (4 * j == 3 * R) {
    return rot<!INV>(a);
  } else {
    constexpr double c = tw<R>::re[j];
    constexpr double s = INV ? -tw<R>::im[j] : tw<R>::im[j];
    return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// a * w for a forward twiddle w taken from a table (conjugated for INV)
template <bool INV>
__device__ __forceinline__ double2 mul_tw(double2 a, double2 w) {
  return INV ? make_double2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y)
             : make_double2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

template <int R>
struct dft_split;
template <> struct dft_split<6> { static constexpr int P = 2, Q = 3; };
template <> struct dft_split<9> { static constexpr int P = 3, Q = 3; };
template <> struct dft_split<12> { static constexpr int P = 4, Q = 3; };
template <> struct dft_split<16> { static constexpr int P = 4, Q = 4; };
template <> struct dft_split<24> { static constexpr int P = 8, Q = 3; };
template <> struct dft_split<32> { static constexpr int P = 8, Q = 4; };

// in-place DFT of v[0..R), natural order in and out
template <bool INV, int R>
__device__ __forceinline__ void dft(double2* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    bfly2<INV>(v);
  } else if constexpr (R == 3) {
    bfly3<INV>(v);
  } else if constexpr (R == 4) {
    bfly4<INV>(v);
  } else if constexpr (R == 5) {
    bfly5<INV>(v);
  } else if constexpr (R == 8) {
    bfly8<INV>(v);
  } else {
    constexpr int P = dft_split<R>::P, Q = dft_split<R>::Q;
    double2 t[R];
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      double2 u[P];
      sfor<0, P>([&](auto pc) { u[decltype(pc)::value] = v[Q * decltype(pc)::value + q]; });
      dft<INV, P>(u);
      sfor<0, P>([&](auto kc) {
        constexpr int k1 = decltype(kc)::value;
        t[k1 * Q + q] = twc<INV, R, q * k1>(u[k1]);
      });
    });
    sfor<0, P>([&](auto kc) {
      constexpr int k1 = decltype(kc)::value;
      double2 u[Q];
      sfor<0, Q>([&](auto qc) { u[decltype(qc)::value] = t[k1 * Q + decltype(qc)::value]; });
      dft<INV, Q>(u);
      sfor<0, Q>([&](auto c) { v[k1 + P * decltype(c)::value] = u[decltype(c)::value]; });
    });
  }
}

}  // namespace cfd
