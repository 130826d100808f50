// fft.hpp -- batched mixed-radix (2,3,4,5) Stockham FFT in LDS for one
// workgroup.  B transforms of length L are stored back to back in LDS; every
// pass reads one buffer and writes the other (autosort), so the result ends in
// natural order without a bit-reversal step.  Twiddles come from an LDS copy
// of tw[m] = exp(-2 pi i m / L).
//
// Sign convention (cft_buff.f:9-13, rft_buff.f:26-47): forward e^{-i},
// inverse e^{+i}; both unnormalised here, scaling is applied by the caller.
#pragma once
#include "common.hpp"

namespace cfd {

template <bool INV>
__device__ __forceinline__ double2 twid(const double2* tw, int m) {
  const double2 w = tw[m];
  return INV ? make_double2(w.x, -w.y) : w;
}

// (-i)*b for the forward transform, (+i)*b for the inverse
template <bool INV>
__device__ __forceinline__ double2 rot(double2 b) {
  return INV ? make_double2(-b.y, b.x) : make_double2(b.y, -b.x);
}

template <bool INV>
__device__ __forceinline__ void bfly2(double2* v) {
  const double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void bfly4(double2* v) {
  const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  const double2 b0 = cadd(v[1], v[3]), b1 = rot<INV>(csub(v[1], v[3]));
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1);
  v[3] = csub(a1, b1);
}

template <bool INV>
__device__ __forceinline__ void bfly3(double2* v) {
  const double s = 0.86602540378443864676;  // sin(2 pi/3)
  const double2 t = cadd(v[1], v[2]);
  const double2 m = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
  const double2 d = rot<INV>(cscl(csub(v[1], v[2]), s));
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, d);
  v[2] = csub(m, d);
}

template <bool INV>
__device__ __forceinline__ void bfly5(double2* v) {
  const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos(2pi/5), cos(4pi/5)
  const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin(2pi/5), sin(4pi/5)
  const double2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  const double2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  const double2 m1 = make_double2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
  const double2 m2 = make_double2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
  const double2 n1 = rot<INV>(make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  const double2 n2 = rot<INV>(make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = make_double2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
  v[1] = cadd(m1, n1);
  v[4] = csub(m1, n1);
  v[2] = cadd(m2, n2);
  v[3] = csub(m2, n2);
}

// One workgroup transforms B sequences of length P.L held in a[t*L + n].
// b is scratch of the same size.  Returns the buffer holding the result.
// Must be called by every thread of the workgroup (it synchronises).
template <bool INV>
__device__ double2* fft_lds(double2* a, double2* b, const double2* tw, const FftPlan& P, int B) {
  const int L = P.L;
  const int tid = threadIdx.x, nth = blockDim.x;
  int Ns = 1;
  for (int s = 0; s < P.nr; ++s) {
    const int R = P.rad[s];
    const int nb = L / R;
    const int ts = L / (Ns * R);
    for (int w = tid; w < B * nb; w += nth) {
      const int t = w / nb, j = w - t * nb;
      const double2* src = a + t * L;
      double2* dst = b + t * L;
      const int k = j % Ns;
      double2 v[5];
      for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
      if (k) {
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<INV>(tw, r * k * ts));
      }
      switch (R) {
        case 2: bfly2<INV>(v); break;
        case 3: bfly3<INV>(v); break;
        case 4: bfly4<INV>(v); break;
        default: bfly5<INV>(v); break;
      }
      const int base = (j - k) * R + k;
      for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
    __syncthreads();
    double2* tmp = a;
    a = b;
    b = tmp;
    Ns *= R;
  }
  return a;
}

__device__ inline void load_twiddles(double2* lds, const double2* g, int L) {
  for (int m = threadIdx.x; m < L; m += blockDim.x) lds[m] = g[m];
}

}  // namespace cfd

// ---------------------------------------------------------------------------
// Compile-time plans.  A transform length known at compile time gets its
// radix sequence (8s first, then 4, 2, 3, 5) folded into the code: constant
// strides and divisors, unrolled butterfly loops, and an LDS layout padded by
// one element in eight so the stride-R stores of the early Stockham passes
// spread over all banks.
// ---------------------------------------------------------------------------
namespace cfd {

struct CPlan {
  int nr;
  int rad[kMaxRadix];
};

constexpr CPlan make_cplan(int L) {
  CPlan p{0, {}};
  int m = L;
  while (m % 8 == 0 && m > 1) { p.rad[p.nr++] = 8; m /= 8; }
  while (m % 4 == 0 && m > 1) { p.rad[p.nr++] = 4; m /= 4; }
  while (m % 2 == 0 && m > 1) { p.rad[p.nr++] = 2; m /= 2; }
  while (m % 3 == 0 && m > 1) { p.rad[p.nr++] = 3; m /= 3; }
  while (m % 5 == 0 && m > 1) { p.rad[p.nr++] = 5; m /= 5; }
  if (m != 1) p.nr = -1;
  return p;
}

// padded LDS index
__device__ __forceinline__ constexpr int pidx(int i) { return i + (i >> 3); }
constexpr int padded(int n) { return n + (n >> 3) + 1; }

template <bool INV>
__device__ __forceinline__ void bfly8(double2* v) {
  // radix-8 as 2 x radix-4 with the internal twiddles w8^k
  const double s = 0.70710678118654752440;
  double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<INV>(e);
  bfly4<INV>(o);
  // o[k] *= w8^k  (w8 = e^{-i pi/4} forward, e^{+i pi/4} inverse)
  o[1] = INV ? make_double2(s * (o[1].x - o[1].y), s * (o[1].x + o[1].y))
             : make_double2(s * (o[1].x + o[1].y), s * (o[1].y - o[1].x));
  o[2] = rot<INV>(o[2]);
  o[3] = INV ? make_double2(-s * (o[3].x + o[3].y), s * (o[3].x - o[3].y))
             : make_double2(s * (o[3].y - o[3].x), -s * (o[3].x + o[3].y));
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

template <bool INV, int R>
__device__ __forceinline__ void bfly(double2* v) {
  if constexpr (R == 2) bfly2<INV>(v);
  else if constexpr (R == 3) bfly3<INV>(v);
  else if constexpr (R == 4) bfly4<INV>(v);
  else if constexpr (R == 5) bfly5<INV>(v);
  else bfly8<INV>(v);
}

template <bool INV, int L, int B, int S, int Ns>
__device__ __forceinline__ void fft_pass(double2*& a, double2*& b, const double2* tw) {
  constexpr CPlan P = make_cplan(L);
  if constexpr (S < P.nr) {
    constexpr int R = P.rad[S];
    constexpr int nb = L / R;
    constexpr int ts = L / (Ns * R);
    for (int w = threadIdx.x; w < B * nb; w += blockDim.x) {
      const int t = w / nb, j = w - t * nb;
      const double2* src = a + t * padded(L);
      double2* dst = b + t * padded(L);
      const int k = j % Ns;
      double2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = src[pidx(j + r * nb)];
      if (Ns > 1 && k) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<INV>(tw, r * k * ts));
      }
      bfly<INV, R>(v);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[pidx(base + r * Ns)] = v[r];
    }
    __syncthreads();
    double2* tmp = a;
    a = b;
    b = tmp;
    fft_pass<INV, L, B, S + 1, Ns * R>(a, b, tw);
  }
}

// B transforms of compile-time length L in padded LDS rows a[t*padded(L) +
// pidx(n)]; b is scratch.  Returns the buffer holding the result.
template <bool INV, int L, int B>
__device__ double2* fft_lds_c(double2* a, double2* b, const double2* tw) {
  static_assert(make_cplan(L).nr > 0, "length must factor into 2, 3, 5");
  fft_pass<INV, L, B, 0, 1>(a, b, tw);
  return a;
}

}  // namespace cfd
