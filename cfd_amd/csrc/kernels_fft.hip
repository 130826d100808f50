// kernels_fft.hip -- the three FFT passes of the nonlinear term and of the
// immersed-boundary transforms.
//
//  z-inverse  spectral plane (N2 x Mx) -> Z plane (Ngz x Mx): complex inverse
//             transform along z of the Mx dealiased x-modes, zero-padded from
//             N2 to Ngz modes (four_to_phys_u first half, FOU3D.f90:546-564).
//             The y interpolations interp_u / interp_v (FOU3D.f90:261-308) are
//             folded into the load.
//  x-pass     per z-row: complex-to-real inverse along x, products, real-to-
//             complex forward along x, without writing physical fields to HBM
//             (ops_in_planes, FOU3D.f90:1529-1617).  Two real rows share one
//             complex transform of length Ngx: a + i b.
//  z-forward  Z plane -> spectral plane, keeping the N2 dealiased z-modes and
//             applying the x/z derivatives (der_x/der_z, FOU3D.f90:310-414).
#include <cstring>

#include "dft.hpp"
#include "fft.hpp"

namespace cfd {

namespace {

constexpr int kThreads = 256;

__device__ inline double dabs(double x) { return x < 0 ? -x : x; }

// slot of reference mode k (0-based) in the Ngz-point z transform
__device__ __forceinline__ int zslot(int k, int N2, int ngz) { return k < N2 / 2 ? k : k - N2 + ngz; }

// LDS row layout: padded rows for compile-time lengths (LC > 0), dense rows
// for the runtime-plan fallback (LC = 0)
template <int LC>
__device__ __forceinline__ int rstride(int L) {
  if constexpr (LC > 0) return padded(LC);
  else return L;
}
template <int LC>
__device__ __forceinline__ int ix(int i) {
  if constexpr (LC > 0) return pidx(i);
  else return i;
}
template <bool INV, int LC, int B>
__device__ __forceinline__ double2* run_fft(double2* a, double2* b, const double2* tw, const FftPlan& P) {
  if constexpr (LC > 0) return fft_lds_c<INV, LC, B>(a, b, tw);
  else return fft_lds<INV>(a, b, tw, P, B);
}

// transforms per block of the z passes
constexpr int ti_for(int L) { return L <= 128 ? 8 : (L <= 256 ? 4 : 2); }
constexpr int kTIFallback = 4;

template <int LC>
size_t lds_bytes(int L, int B) {
  const int st = LC > 0 ? padded(LC) : L;
  return sizeof(double2) * (size_t)(L + 2 * B * st);
}

// ---------------------------------------------------------------------------
// z-inverse: grid (ceil(Mx/TI), njobs)
// ---------------------------------------------------------------------------
template <int LC, int TI>
__global__ void __launch_bounds__(kThreads) k_zinv(const ZInvJob* __restrict__ jobs, FftPlan P,
                                                   const double2* __restrict__ twg, int N2, int Mx) {
  extern __shared__ double2 lds[];
  const int L = LC > 0 ? LC : P.L;
  const int st = rstride<LC>(L);
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + TI * st;
  const ZInvJob J = jobs[blockIdx.y];
  const int i0 = blockIdx.x * TI;
  const int ncol = N2 * Mx;
  load_twiddles(tw, twg, L);
  // zero the padding slots, then scatter the N2 dealiased modes
  for (int w = threadIdx.x; w < TI * st; w += blockDim.x) A[w] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int w = threadIdx.x; w < N2 * TI; w += blockDim.x) {
    const int k = w / TI, t = w - k * TI;
    const int i = i0 + t;
    if (i >= Mx) continue;
    const size_t off = (size_t)k * Mx + i;
    double2 v;
    if ((J.flags & 1) && k == N2 / 2) {
      v = make_double2(0.0, 0.0);
    } else if (J.mode == 0) {
      v = J.src[(size_t)J.pa * ncol + off];
    } else {
      const double2 a = J.src[(size_t)J.pa * ncol + off], b = J.src[(size_t)J.pb * ncol + off];
      v = make_double2(J.wa * a.x + J.wb * b.x, J.wa * a.y + J.wb * b.y);
      if (J.mode == 2) {
        const double2 c = J.src[(size_t)J.pc * ncol + off];
        v = make_double2(J.g * (c.x - v.x) + v.x, J.g * (c.y - v.y) + v.y);
      }
    }
    A[t * st + ix<LC>(zslot(k, N2, L))] = v;
  }
  __syncthreads();
  double2* R = run_fft<true, LC, TI>(A, B, tw, P);
  for (int w = threadIdx.x; w < L * TI; w += blockDim.x) {
    const int z = w / TI, t = w - z * TI;
    const int i = i0 + t;
    if (i >= Mx) continue;
    ((double2*)J.dst)[((size_t)J.pd * L + z) * Mx + i] = R[t * st + ix<LC>(z)];
  }
}

// ---------------------------------------------------------------------------
// x-pass helpers
// ---------------------------------------------------------------------------
// Hermitian pair spectrum of a + i b from the Mx modes of A and B; the
// imaginary parts of the k=0 terms are ignored as rfftb ignores them
// (rft_buff.f:78-83).
__device__ __forceinline__ double2 pair_mode(const double2* __restrict__ A, const double2* __restrict__ B, int m,
                                             int L, int Mx) {
  if (m == 0) {
    const double a = A ? A[0].x : 0.0, b = B ? B[0].x : 0.0;
    return make_double2(a, b);
  }
  if (m < Mx) {
    const double2 a = A ? A[m] : make_double2(0, 0), b = B ? B[m] : make_double2(0, 0);
    return make_double2(a.x - b.y, a.y + b.x);
  }
  if (m > L - Mx) {
    const int q = L - m;
    const double2 a = A ? A[q] : make_double2(0, 0), b = B ? B[q] : make_double2(0, 0);
    return make_double2(a.x + b.y, -a.y + b.x);
  }
  return make_double2(0.0, 0.0);
}

// split the forward transform F of a + i b into the modes of a and b, scaled
template <int LC>
__device__ __forceinline__ void split_mode(const double2* F, int m, int L, double sc, double2& a, double2& b) {
  const double2 f = F[ix<LC>(m)], g = F[ix<LC>(m == 0 ? 0 : L - m)];
  a = make_double2(0.5 * sc * (f.x + g.x), 0.5 * sc * (f.y - g.y));
  b = make_double2(0.5 * sc * (f.y + g.y), -0.5 * sc * (f.x - g.x));
}

__device__ inline void atomic_max_pos(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = fmax(r, red[q]);
  return r;
}

// ---------------------------------------------------------------------------
// nonlinear x-pass: grid (Ngz, nyu).  Block (z, q) handles u-plane q and, when
// q < nyv, v-plane q of row z.
// Z[0..2]: u1, u2_itp, u3 (u planes)   Z[3..5]: u1_itp, u2, u3_itp (v planes)
// Z[6..9]: uu, uw, ww, vv (u planes)   Z[10..11]: uv, vw (v planes)
// ---------------------------------------------------------------------------
struct XNLArgs {
  const void* zi[6];  // double2 (FP64) or float2 (mixed) Z planes
  void* zo[6];
  const double* dthdyv;
  DevScalars* sc;
  int nyu, nyv, Mx, ngz, cfl;
  int zs;   // Z row stride (elements)
  int top;  // also the products of the top ghost u plane (the last-plane replay reads them)
  // per u / v plane: 1 when the plane's maxima count in this solver's CFL
  // (multi-band resolution: each resolution takes its own planes); null: all
  const unsigned char *cfu, *cfv;
  // plane window {lo, hi, s0, s1}: the grid covers planes [lo, hi) without
  // [s0, s1) (multi-band resolution: each resolution transforms its own
  // planes); {0, nyu, 0, 0} covers all
  int win[4];
  // a zeroed row of at least Mx elements, read in place of an absent row
  // (a z beyond Ngz, the v-plane pairs of the last u plane), so that every
  // load of the four-step x-pass is unconditional
  const void* zrow;
};

// plane of y-block b of the window {lo, hi, s0, s1}
__device__ __forceinline__ int win_plane(int b, const int* w) {
  const int q = w[0] + b;
  return q < w[2] ? q : q + (w[3] - w[2]);
}
inline int win_count(const int* w) { return (w[1] - w[0]) - (w[3] - w[2]); }

template <int LC>
__global__ void __launch_bounds__(kThreads) k_xpass_nl(XNLArgs a, FftPlan P, const double2* __restrict__ twg) {
  extern __shared__ double2 lds[];
  const int L = LC > 0 ? LC : P.L, Mx = a.Mx;
  const int st = rstride<LC>(L);
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + 3 * st;
  __shared__ double red[kThreads / 64];
  const int z = blockIdx.x, q = win_plane(blockIdx.y, a.win);
  const int ngz = gridDim.x;
  const bool hasv = q < a.nyv;
  const bool uint = q > 0 && (q < a.nyu - 1 || (a.top && q == a.nyu - 1));
  const bool vint = hasv && q > 0 && q < a.nyv - 1;
  load_twiddles(tw, twg, L);
  const size_t ru = ((size_t)q * ngz + z) * Mx;
  const double2* u1 = (const double2*)a.zi[0] + ru;
  const double2* u2i = (const double2*)a.zi[1] + ru;
  const double2* u3 = (const double2*)a.zi[2] + ru;
  const double2* u1i = hasv ? (const double2*)a.zi[3] + ru : nullptr;
  const double2* v = hasv ? (const double2*)a.zi[4] + ru : nullptr;
  const double2* u3i = hasv ? (const double2*)a.zi[5] + ru : nullptr;
  for (int m = threadIdx.x; m < L; m += blockDim.x) {
    A[ix<LC>(m)] = pair_mode(u1, u3, m, L, Mx);
    A[st + ix<LC>(m)] = pair_mode(u2i, v, m, L, Mx);
    A[2 * st + ix<LC>(m)] = pair_mode(u1i, u3i, m, L, Mx);
  }
  __syncthreads();
  double2* X = run_fft<true, LC, 3>(A, B, tw, P);
  double2* Y = (X == A) ? B : A;
  // products in physical space; CFL maxima (dtc_calc, FOU3D.f90:516-544)
  const double dv = vint ? a.dthdyv[q] : 0.0;
  double m1 = 0.0, m2 = 0.0, m3 = 0.0;
  for (int x = threadIdx.x; x < L; x += blockDim.x) {
    const int o = ix<LC>(x);
    const double2 p0 = X[o], p1 = X[st + o], p2 = X[2 * st + o];
    const double uu = p0.x, ww = p0.y, vc = p1.x, vf = p1.y, uf = p2.x, wf = p2.y;
    Y[o] = make_double2(uu * uu, uu * ww);
    Y[st + o] = make_double2(vc * vc, ww * ww);
    Y[2 * st + o] = make_double2(uf * vf, wf * vf);
    m1 = fmax(m1, dabs(uu));
    m3 = fmax(m3, dabs(ww));
    m2 = fmax(m2, dabs(vf) * dv);
  }
  if (a.cfu && !a.cfu[q]) m1 = m3 = 0.0;
  if (a.cfv && !(hasv && a.cfv[q])) m2 = 0.0;
  if (a.cfl) {
    m1 = block_max(m1, red);
    if (threadIdx.x == 0) atomic_max_pos(&a.sc->m1, m1);
    m3 = block_max(m3, red);
    if (threadIdx.x == 0) atomic_max_pos(&a.sc->m3, m3);
    m2 = block_max(m2, red);
    if (threadIdx.x == 0) atomic_max_pos(&a.sc->m2, m2);
  }
  __syncthreads();
  const double2* F = run_fft<false, LC, 3>(Y, X, tw, P);
  const double sc = 1.0 / L;
  for (int m = threadIdx.x; m < Mx; m += blockDim.x) {
    double2 ca, cb;
    if (uint) {
      split_mode<LC>(F, m, L, sc, ca, cb);
      ((double2*)a.zo[0])[ru + m] = ca;  // uu
      ((double2*)a.zo[1])[ru + m] = cb;  // uw
      split_mode<LC>(F + st, m, L, sc, ca, cb);
      ((double2*)a.zo[3])[ru + m] = ca;  // vv
      ((double2*)a.zo[2])[ru + m] = cb;  // ww
    }
    if (hasv) {
      split_mode<LC>(F + 2 * st, m, L, sc, ca, cb);
      ((double2*)a.zo[4])[ru + m] = ca;  // uv
      ((double2*)a.zo[5])[ru + m] = cb;  // vw
    }
  }
}

// ---------------------------------------------------------------------------
// z-forward of the nonlinear products with the x/z derivatives.
// grid (ceil(Mx/TI), nyu + nyv): y-block q < nyu is u-plane q, else v-plane.
// ---------------------------------------------------------------------------
struct ZNLArgs {
  const void* z[6];  // uu uw ww vv (u planes), uv vw (v planes); double2 or float2
  void *nxz1, *nxz3, *fvv, *nxz2, *fuv, *fvw;
  const double *k1x, *k1z;
  int nyu, nyv, N2, Mx;
  int zs;   // Z row stride (elements)
  int top;  // also the x/z terms of the top ghost u plane and the top wall v plane
  // plane windows {lo, hi, s0, s1} of the u and v planes transformed
  // (multi-band resolution: each resolution transforms its own planes)
  int winu[4], winv[4];
};

constexpr int kMaxOut = 8;  // outputs (k,t) per thread: N2*TI <= kMaxOut*kThreads

template <int LC, int TI>
__global__ void __launch_bounds__(kThreads) k_zfwd_nl(ZNLArgs a, FftPlan P, const double2* __restrict__ twg) {
  extern __shared__ double2 lds[];
  const int L = LC > 0 ? LC : P.L, Mx = a.Mx, N2 = a.N2;
  const int st = rstride<LC>(L);
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + TI * st;
  const int i0 = blockIdx.x * TI;
  const int nue = (a.winu[1] - a.winu[0]) - (a.winu[3] - a.winu[2]);
  const bool isu = (int)blockIdx.y < nue;
  const int q = isu ? win_plane(blockIdx.y, a.winu) : win_plane(blockIdx.y - nue, a.winv);
  const int ny = isu ? a.nyu : a.nyv;
  const bool inter = q > 0 && (q < ny - 1 || (a.top && q == ny - 1));
  if (isu && !inter) return;  // ghost u planes: no product is needed there (but see a.top)
  load_twiddles(tw, twg, L);
  const int ncol = N2 * Mx;
  const double sc = 1.0 / L;
  double2 acc1[kMaxOut], acc2[kMaxOut];
#pragma unroll
  for (int c = 0; c < kMaxOut; ++c) acc1[c] = acc2[c] = make_double2(0, 0);
  const int nprod = isu ? 4 : 2;
  for (int pr = 0; pr < nprod; ++pr) {
    const double2* src = (const double2*)a.z[isu ? pr : 4 + pr];
    for (int w = threadIdx.x; w < TI * L; w += blockDim.x) {
      const int zz = w / TI, t = w - zz * TI;
      const int i = i0 + t;
      A[t * st + ix<LC>(zz)] = (i < Mx) ? src[((size_t)q * L + zz) * Mx + i] : make_double2(0, 0);
    }
    __syncthreads();
    const double2* R = run_fft<false, LC, TI>(A, B, tw, P);
#pragma unroll
    for (int c = 0; c < kMaxOut; ++c) {
      const int w = threadIdx.x + c * blockDim.x;
      if (w >= N2 * TI) break;
      const int k = w / TI, t = w - k * TI;
      const int i = i0 + t;
      if (i >= Mx) continue;
      const int sl = zslot(k, N2, L);
      const double2 X = cscl(R[t * st + ix<LC>(sl)], sc);
      const bool zero = (sl == L - N2 / 2);  // der_x/der_z zero band (FOU3D.f90:329-333)
      const double kx = a.k1x[i], kz = a.k1z[k];
      const size_t o = (size_t)q * ncol + (size_t)k * Mx + i;
      if (isu) {
        // Nu1 = ikx uu + ikz uw ; Nu3 = ikx uw + ikz ww ; keep vv
        if (pr == 0) acc1[c] = imulk(kx, X);
        if (pr == 1) { acc1[c] = cadd(acc1[c], imulk(kz, X)); acc2[c] = imulk(kx, X); }
        if (pr == 2) acc2[c] = cadd(acc2[c], imulk(kz, X));
        if (pr == 3) {
          ((double2*)a.fvv)[o] = X;
          ((double2*)a.nxz1)[o] = zero ? make_double2(0, 0) : acc1[c];
          ((double2*)a.nxz3)[o] = zero ? make_double2(0, 0) : acc2[c];
        }
      } else {
        // Nu2 = ikx uv + ikz vw ; keep uv, vw
        if (pr == 0) { acc1[c] = imulk(kx, X); ((double2*)a.fuv)[o] = X; }
        if (pr == 1) {
          ((double2*)a.fvw)[o] = X;
          if (inter) ((double2*)a.nxz2)[o] = zero ? make_double2(0, 0) : cadd(acc1[c], imulk(kz, X));
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// IB transforms
// ---------------------------------------------------------------------------
// inverse pair: spectral rows (a, b) of one plane -> physical rows; grid (Ngz, njobs)
template <int LC>
__global__ void __launch_bounds__(kThreads) k_xpass_inv2(const XJob* __restrict__ jobs, FftPlan P,
                                                         const double2* __restrict__ twg, int Mx) {
  extern __shared__ double2 lds[];
  const int L = LC > 0 ? LC : P.L;
  const int st = rstride<LC>(L);
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + st;
  const XJob J = jobs[blockIdx.y];
  const int z = blockIdx.x, ngz = gridDim.x;
  load_twiddles(tw, twg, L);
  const size_t r = ((size_t)J.plane * ngz + z) * Mx;
  for (int m = threadIdx.x; m < L; m += blockDim.x) A[ix<LC>(m)] = pair_mode(J.za + r, J.zb + r, m, L, Mx);
  __syncthreads();
  const double2* X = run_fft<true, LC, 1>(A, B, tw, P);
  const size_t o = ((size_t)J.plane * ngz + z) * L;
  for (int x = threadIdx.x; x < L; x += blockDim.x) {
    const double2 v = X[ix<LC>(x)];
    J.pa[o + x] = v.x;
    J.pb[o + x] = v.y;
  }
}

// forward: physical rows z, z+1 of one plane -> spectral rows; grid (Ngz/2, njobs)
template <int LC>
__global__ void __launch_bounds__(kThreads) k_xpass_fwd2(const XJob* __restrict__ jobs, FftPlan P,
                                                         const double2* __restrict__ twg, int Mx) {
  extern __shared__ double2 lds[];
  const int L = LC > 0 ? LC : P.L;
  const int st = rstride<LC>(L);
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + st;
  const XJob J = jobs[blockIdx.y];
  const int z = 2 * blockIdx.x, ngz = 2 * gridDim.x;
  load_twiddles(tw, twg, L);
  const size_t o = ((size_t)J.plane * ngz + z) * L;
  for (int x = threadIdx.x; x < L; x += blockDim.x) A[ix<LC>(x)] = make_double2(J.pa[o + x], J.pa[o + L + x]);
  __syncthreads();
  const double2* F = run_fft<false, LC, 1>(A, B, tw, P);
  const size_t r = ((size_t)J.plane * ngz + z) * Mx;
  const double sc = 1.0 / L;
  for (int m = threadIdx.x; m < Mx; m += blockDim.x) {
    double2 ca, cb;
    split_mode<LC>(F, m, L, sc, ca, cb);
    J.zo[r + m] = ca;
    J.zo[r + Mx + m] = cb;
  }
}

// plain z-forward: Z plane -> N2 dealiased spectral modes; grid (ceil(Mx/TI), njobs)
template <int LC, int TI>
__global__ void __launch_bounds__(kThreads) k_zfwd_plain(const ZFwdJob* __restrict__ jobs, FftPlan P,
                                                         const double2* __restrict__ twg, int N2, int Mx) {
  extern __shared__ double2 lds[];
  const int L = LC > 0 ? LC : P.L;
  const int st = rstride<LC>(L);
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + TI * st;
  const ZFwdJob J = jobs[blockIdx.y];
  const int i0 = blockIdx.x * TI;
  load_twiddles(tw, twg, L);
  for (int w = threadIdx.x; w < TI * L; w += blockDim.x) {
    const int zz = w / TI, t = w - zz * TI;
    const int i = i0 + t;
    A[t * st + ix<LC>(zz)] = (i < Mx) ? J.src[((size_t)J.ps * L + zz) * Mx + i] : make_double2(0, 0);
  }
  __syncthreads();
  const double2* R = run_fft<false, LC, TI>(A, B, tw, P);
  const double sc = 1.0 / L;
  const int ncol = N2 * Mx;
  for (int w = threadIdx.x; w < N2 * TI; w += blockDim.x) {
    const int k = w / TI, t = w - k * TI;
    const int i = i0 + t;
    if (i >= Mx) continue;
    J.dst[(size_t)J.pd * ncol + (size_t)k * Mx + i] = cscl(R[t * st + ix<LC>(zslot(k, N2, L))], sc);
  }
}

// ---------------------------------------------------------------------------
// Four-step kernels.  A length L = P x Q transform is done as P- and Q-point
// register DFTs (dft.hpp) with one exchange through LDS in between: input
// n = Q p + q, output k = k1 + P k2 (the identity of the dft.hpp header applied
// at length L, with the L-point twiddles w_L^{q k1} read from the table).
// The z kernels load and store HBM straight from registers, TI x-modes per
// block side by side (lanes of consecutive i touch one TI-element segment).
// Sizes follow the dealiasing rule of the reference: N = 2*(L/3) modes kept
// (N2 along z, N1 along x with Mx = N1/2 + 1), FOU3D.f90:546-564.
// T is the arithmetic and storage type of the Z planes and of the nonlinear
// outputs: double for the FP64 path, float for the mixed path (whose
// spectral velocity inputs stay FP64 and are rounded once on load).
// ---------------------------------------------------------------------------
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int rup64(int n) { return (n + 63) / 64 * 64; }

template <int L>
struct Dealias {
  static constexpr int N = 2 * (L / 3);  // kept modes
  static constexpr int H = N / 2;        // slots [H, L-H) carry no mode
  static constexpr int Mx = N / 2 + 1;   // kept x-modes of a real row
};

template <typename C>
__device__ __forceinline__ C from_d2(double2 v) {
  return mk<C>((real_t<C>)v.x, (real_t<C>)v.y);
}

// z-inverse, four-step: TI*max(P,Q) threads; LDS X[k1][q][t].  Destination
// rows have stride zs elements.  Block -> (job, x-tile): with jper = 0 the
// grid is (ntiles, njobs); with jper > 0 it is one-dimensional and XCD-aware:
// blocks are dealt to the 8 XCDs round-robin by block id, so block id b runs
// on XCD b % 8, and XCD g takes the contiguous job range [g jper, (g+1) jper).
// With the jobs ordered by y-plane, the jobs that read the same source
// planes (a plane and its interpolation neighbours) then share one L2.
// Block (bx, by) of the grid; the kernel k_zinv4 runs it for one solver,
// k_zinv4_b for the members of a batch.
template <typename T, int L, int P, int Q, int TI>
__device__ __forceinline__ void zinv4_body(const ZInvJob* __restrict__ jobs, const cplx_t<T>* __restrict__ twg,
                                           int Mx, int zs, int njobs, int ntiles, int jper, int bx, int by) {
  using C = cplx_t<T>;
  static_assert(P * Q == L, "P*Q");
  constexpr int N2 = Dealias<L>::N, H = Dealias<L>::H;
  __shared__ C X[P * Q * TI];
  int job, tile;
  if (jper > 0) {
    const int g = bx & 7, kk = bx >> 3;
    const int jl = kk / ntiles;
    job = g * jper + jl;
    tile = kk - jl * ntiles;
    if (jl >= jper || job >= njobs) return;
  } else {
    job = by;
    tile = bx;
  }
  const ZInvJob J = jobs[job];
  const int tid = threadIdx.x, i0 = tile * TI;
  const size_t ncol = (size_t)N2 * Mx;
  if (tid < Q * TI) {
    const int t = tid % TI, q = tid / TI, i = i0 + t;
    // element offsets of the P inputs (-1: zero); all loads of the thread are
    // issued before any of them is used (the job's mode is uniform)
    int o[P];
    sfor<0, P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      if constexpr (Q * p >= H && Q * p + Q <= L - H) {
        o[p] = -1;  // the whole group lies in the empty band
      } else {
        const int n = Q * p + q;
        int k = n < H ? n : (n >= L - H ? n - L + N2 : -1);
        if ((J.flags & 1) && k == H) k = -1;
        o[p] = (k >= 0 && i < Mx) ? k * Mx + i : -1;
      }
    });
    // global-address-space loads of 16-byte vectors (the job's pointer,
    // read from memory, would otherwise be a flat pointer)
    typedef double d2v __attribute__((ext_vector_type(2)));
    using G = const __attribute__((address_space(1))) d2v*;
    auto ld = [&](G src, int off) {
      if (off < 0) return make_double2(0.0, 0.0);
      const d2v r = src[off];
      return make_double2(r.x, r.y);
    };
    const G srcA = (G)(J.src + (size_t)J.pa * ncol);
    double2 x[P];
    sfor<0, P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      x[p] = ld(srcA, o[p]);
    });
    if (J.mode != 0) {
      // y interpolation (interp_u / interp_v); mode 2 adds the ghost weight
      const G srcB = (G)(J.src + (size_t)J.pb * ncol);
      double2 y[P];
      sfor<0, P>([&](auto pc) {
        constexpr int p = decltype(pc)::value;
        y[p] = ld(srcB, o[p]);
      });
      if (J.mode == 2) {
        const G srcC = (G)(J.src + (size_t)J.pc * ncol);
        double2 z[P];
        sfor<0, P>([&](auto pc) {
          constexpr int p = decltype(pc)::value;
          z[p] = ld(srcC, o[p]);
        });
        sfor<0, P>([&](auto pc) {
          constexpr int p = decltype(pc)::value;
          const double2 w = make_double2(J.wa * x[p].x + J.wb * y[p].x, J.wa * x[p].y + J.wb * y[p].y);
          x[p] = make_double2(J.g * (z[p].x - w.x) + w.x, J.g * (z[p].y - w.y) + w.y);
        });
      } else {
        sfor<0, P>([&](auto pc) {
          constexpr int p = decltype(pc)::value;
          x[p] = make_double2(J.wa * x[p].x + J.wb * y[p].x, J.wa * x[p].y + J.wb * y[p].y);
        });
      }
    }
    C v[P];
    sfor<0, P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      v[p] = from_d2<C>(x[p]);
    });
    dft<true, P>(v);
    X[q * TI + t] = v[0];
    sfor<1, P>([&](auto kc) {
      constexpr int k1 = decltype(kc)::value;
      X[(k1 * Q + q) * TI + t] = mul_tw<true>(v[k1], twg[q * k1]);
    });
  }
  __syncthreads();
  if (tid < P * TI) {
    const int t = tid % TI, k1 = tid / TI, i = i0 + t;
    C u[Q];
    sfor<0, Q>([&](auto qc) { u[decltype(qc)::value] = X[(k1 * Q + decltype(qc)::value) * TI + t]; });
    dft<true, Q>(u);
    if (i < Mx) {
      C* d = (C*)J.dst + (size_t)J.pd * L * zs + i;
      sfor<0, Q>([&](auto c) { d[(size_t)(k1 + P * decltype(c)::value) * zs] = u[decltype(c)::value]; });
    }
  }
}

template <typename T, int L, int P, int Q, int TI>
__global__ void __launch_bounds__(TI * cmax(P, Q)) k_zinv4(const ZInvJob* __restrict__ jobs,
                                                          const cplx_t<T>* __restrict__ twg, int Mx, int zs,
                                                          int njobs, int ntiles, int jper) {
  zinv4_body<T, L, P, Q, TI>(jobs, twg, Mx, zs, njobs, ntiles, jper, blockIdx.x, blockIdx.y);
}

// one member's z-inverse in a batch
struct ZInvArgs {
  const ZInvJob* jobs;
  int Mx, zs, njobs, ntiles, jper;
};

template <typename T, int L, int P, int Q, int TI>
__global__ void __launch_bounds__(TI * cmax(P, Q)) k_zinv4_b(const ZInvArgs* __restrict__ as, BMap bm,
                                                            const cplx_t<T>* __restrict__ twg) {
  int bx = 0, by = 0;
  const int m = bmap_member(bm, blockIdx.x, bx, by);
  if (m < 0) return;
  const ZInvArgs a = as[m];
  zinv4_body<T, L, P, Q, TI>(a.jobs, twg, a.Mx, a.zs, a.njobs, a.ntiles, a.jper, bx, by);
}

// (0,k) * u
template <typename C>
__device__ __forceinline__ C imk(real_t<C> k, C u) {
  return mk<C>(-k * u.y, k * u.x);
}

// z-forward of the nonlinear products, four-step: grid (ceil(Mx/TI), nyu+nyv).
// Output slot k1 + P k2; groups of k2 wholly in the empty band are skipped at
// compile time.  Same outputs as k_zfwd_nl.  The products of a plane are
// transformed one after the other, each output written as soon as it is
// complete (u planes: uu, uw -> Nu1 = ikx uu + ikz uw; ww -> Nu3 = ikx uw +
// ikz ww; vv kept; v planes: uv kept, vw kept -> Nu2 = ikx uv + ikz vw), so one
// accumulator per output slot is live at a time.
template <typename T, int L, int P, int Q, int TI, bool ISU, int PR>
__device__ __forceinline__ void zfwd_product(const ZNLArgs& a, const cplx_t<T>* __restrict__ twg, cplx_t<T>* X,
                                             int pl, bool inter, cplx_t<T>* acc, int bx) {
  using C = cplx_t<T>;
  constexpr int N2 = Dealias<L>::N, H = Dealias<L>::H;
  const int Mx = a.Mx;
  const int tid = threadIdx.x, i0 = bx * TI;
  const size_t ncol = (size_t)N2 * Mx;
  const T sc = (T)(1.0 / L);
  const C* src = (const C*)a.z[ISU ? PR : 4 + PR] + (size_t)pl * L * a.zs;
  if (tid < Q * TI) {
    const int t = tid % TI, q = tid / TI, i = i0 + t;
    C v[P];
    sfor<0, P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      v[p] = i < Mx ? src[(size_t)(Q * p + q) * a.zs + i] : mk<C>(0, 0);
    });
    dft<false, P>(v);
    X[q * TI + t] = v[0];
    sfor<1, P>([&](auto kc) {
      constexpr int kk = decltype(kc)::value;
      X[(kk * Q + q) * TI + t] = mul_tw<false>(v[kk], twg[q * kk]);
    });
  }
  __syncthreads();
  const int t2 = tid % TI, k1 = tid / TI, i2 = i0 + t2;
  if (tid < P * TI && i2 < Mx) {
    C u[Q];
    sfor<0, Q>([&](auto qc) { u[decltype(qc)::value] = X[(k1 * Q + decltype(qc)::value) * TI + t2]; });
    dft<false, Q>(u);
    const T kx = (T)a.k1x[i2];
    sfor<0, Q>([&](auto c) {
      constexpr int k2 = decltype(c)::value;
      if constexpr (!(P * k2 >= H && P * k2 + P <= L - H)) {
        const int sl = k1 + P * k2;
        if (sl < H || sl >= L - H) {
          const int k = sl < H ? sl : sl - L + N2;
          const C Xs = mk<C>(u[k2].x * sc, u[k2].y * sc);
          const bool zero = (sl == L - H);  // der_x/der_z zero band (FOU3D.f90:329-333)
          const T kz = (T)a.k1z[k];
          const size_t o = (size_t)pl * ncol + (size_t)k * Mx + i2;
          if constexpr (ISU) {
            if constexpr (PR == 0) {
              acc[k2] = imk(kx, Xs);  // ikx uu
            } else if constexpr (PR == 1) {
              ((C*)a.nxz1)[o] = zero ? mk<C>(0, 0) : vadd(acc[k2], imk(kz, Xs));  // + ikz uw
              acc[k2] = imk(kx, Xs);                                             // ikx uw
            } else if constexpr (PR == 2) {
              ((C*)a.nxz3)[o] = zero ? mk<C>(0, 0) : vadd(acc[k2], imk(kz, Xs));  // + ikz ww
            } else {
              ((C*)a.fvv)[o] = Xs;
            }
          } else {
            if constexpr (PR == 0) {
              acc[k2] = imk(kx, Xs);  // ikx uv
              ((C*)a.fuv)[o] = Xs;
            } else {
              ((C*)a.fvw)[o] = Xs;
              if (inter) ((C*)a.nxz2)[o] = zero ? mk<C>(0, 0) : vadd(acc[k2], imk(kz, Xs));
            }
          }
        }
      }
    });
  }
  __syncthreads();
}

template <typename T, int L, int P, int Q, int TI>
__device__ __forceinline__ void zfwd4_body(const ZNLArgs& a, const cplx_t<T>* __restrict__ twg, int bx, int by) {
  using C = cplx_t<T>;
  static_assert(P * Q == L, "P*Q");
  __shared__ C X[P * Q * TI];
  const int nue = (a.winu[1] - a.winu[0]) - (a.winu[3] - a.winu[2]);
  const bool isu = by < nue;
  const int pl = isu ? win_plane(by, a.winu) : win_plane(by - nue, a.winv);
  const int ny = isu ? a.nyu : a.nyv;
  const bool inter = pl > 0 && (pl < ny - 1 || (a.top && pl == ny - 1));
  if (isu && !inter) return;  // ghost u planes: no product is needed there (but see a.top)
  C acc[Q];
  if (isu) {
    zfwd_product<T, L, P, Q, TI, true, 0>(a, twg, X, pl, inter, acc, bx);
    zfwd_product<T, L, P, Q, TI, true, 1>(a, twg, X, pl, inter, acc, bx);
    zfwd_product<T, L, P, Q, TI, true, 2>(a, twg, X, pl, inter, acc, bx);
    zfwd_product<T, L, P, Q, TI, true, 3>(a, twg, X, pl, inter, acc, bx);
  } else {
    zfwd_product<T, L, P, Q, TI, false, 0>(a, twg, X, pl, inter, acc, bx);
    zfwd_product<T, L, P, Q, TI, false, 1>(a, twg, X, pl, inter, acc, bx);
  }
}

template <typename T, int L, int P, int Q, int TI>
__global__ void __launch_bounds__(TI * cmax(P, Q)) k_zfwd4(ZNLArgs a, const cplx_t<T>* __restrict__ twg) {
  zfwd4_body<T, L, P, Q, TI>(a, twg, blockIdx.x, blockIdx.y);
}

template <typename T, int L, int P, int Q, int TI>
__global__ void __launch_bounds__(TI * cmax(P, Q)) k_zfwd4_b(const ZNLArgs* __restrict__ as, BMap bm,
                                                            const cplx_t<T>* __restrict__ twg) {
  int bx = 0, by = 0;
  const int m = bmap_member(bm, blockIdx.x, bx, by);
  if (m < 0) return;
  const ZNLArgs a = as[m];
  zfwd4_body<T, L, P, Q, TI>(a, twg, bx, by);
}

// one Hermitian pair mode with the range of m known at compile time where
// the whole group of Q consecutive m shares it; A and B are valid rows (a
// zeroed row stands in for an absent one), so the loads are unconditional
template <int L, int Q, int P0, typename C>
__device__ __forceinline__ C pair_mode_q(const C* __restrict__ A, const C* __restrict__ B, int m) {
  constexpr int Mx = Dealias<L>::Mx;
  constexpr int lo = Q * P0, hi = Q * P0 + Q - 1;
  if constexpr (lo >= Mx && hi <= L - Mx) {
    return mk<C>(0, 0);
  } else if constexpr (lo > 0 && hi < Mx) {
    const C a = A[m], b = B[m];
    return mk<C>(a.x - b.y, a.y + b.x);
  } else if constexpr (lo > L - Mx) {
    const C a = A[L - m], b = B[L - m];
    return mk<C>(a.x + b.y, -a.y + b.x);
  } else if constexpr (lo == 0 && hi < Mx) {
    // m = 0 keeps the real parts only (rfftb drops the imaginary ones)
    const C a = A[m], b = B[m];
    return m == 0 ? mk<C>(a.x, b.x) : mk<C>(a.x - b.y, a.y + b.x);
  } else {
    // a group that straddles Mx or L - Mx: a clamped load, then a select
    const bool low = m < Mx, high = m > L - Mx;
    const int i = low ? m : (high ? L - m : 0);
    const C a = A[i], b = B[i];
    const C f = low ? mk<C>(a.x - b.y, a.y + b.x) : mk<C>(a.x + b.y, -a.y + b.x);
    return (low || high) ? f : mk<C>(0, 0);
  }
}

// lanes per transform of the x-pass block: one group of waves per transform,
// one thread per phase-2 task (k1, r); phases 1 and 4 loop over their Q R tasks
constexpr int xp_width(int P, int Q, int R) { return rup64(P * R); }

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// nonlinear x-pass, four-step: grid (ceil(Ngz/R), nyu), 3 W threads with
// W = xp_width(P, Q, R).  Block (zb, q) takes rows z = R zb .. R zb + R-1 of
// u-plane q and, when q < nyv, of v-plane q; three complex transforms per row
// (pairs (u1,u3), (u2i,v), (u1i,u3i)), transform tr handled by threads
// [tr W, tr W + W) so every branch on tr is uniform across a wavefront.  Per
// transform the LDS slots hold the exchange [k1][q] (row stride Q+1), then the
// physical values in place, then the forward exchange [m2][k1] (row stride
// P+1) and the spectrum in place.
template <typename T, int L, int P, int Q, int R>
__device__ __forceinline__ void xpass4_body(const XNLArgs& a, const cplx_t<T>* __restrict__ twg, int bx, int by) {
  using C = cplx_t<T>;
  static_assert(P * Q == L, "P*Q");
  constexpr int Mx = Dealias<L>::Mx;
  constexpr int W = xp_width(P, Q, R);
  constexpr int SP = cmax(P * (Q + 1), Q * (P + 1));  // slots per transform and row
  __shared__ C S[R * 3 * SP];
  const int tid = threadIdx.x;
  const int tr = tid / W, lane = tid - tr * W;
  const int q = win_plane(by, a.win), z0 = bx * R, ngz = a.ngz;
  const bool hasv = q < a.nyv;
  const bool uint = q > 0 && (q < a.nyu - 1 || (a.top && q == a.nyu - 1));
  const bool vint = hasv && q > 0 && q < a.nyv - 1;
  const C* const* zi = reinterpret_cast<const C* const*>(a.zi);

  // phase 1: tasks (qq, r) of transform tr; inverse P-point DFTs over p of
  // c[Q p + qq].  The Q R tasks are dealt to the lanes in NS = ceil(Q R / W)
  // slots; a lane without a task in the last slot repeats the last task and
  // stores nothing.  Every slot's inputs are loaded before any slot's DFT, so
  // all of the loads are in flight together (the LDS, not the registers,
  // bounds this kernel's residency)
  {
    constexpr int NS = (Q * R + W - 1) / W;
    C v[NS][P];
    int qs[NS], rs[NS];
    sfor<0, NS>([&](auto sc) {
      constexpr int t = decltype(sc)::value;
      const int jt = t * W + lane;
      const int j = (t * W + W <= Q * R || jt < Q * R) ? jt : Q * R - 1;
      const int qq = j % Q, r = j / Q;
      qs[t] = qq;
      rs[t] = r;
      const int z = z0 + r;
      const C* const zr = static_cast<const C*>(a.zrow);
      const C *A = zr, *B = zr;
      if (z < ngz) {
        const size_t ru = ((size_t)q * ngz + z) * a.zs;
        if (tr == 0) {
          A = zi[0] + ru;
          B = zi[2] + ru;
        } else if (tr == 1) {
          A = zi[1] + ru;
          if (hasv) B = zi[4] + ru;
        } else if (hasv) {
          A = zi[3] + ru;
          B = zi[5] + ru;
        }
      }
      sfor<0, P>([&](auto pc) {
        constexpr int p = decltype(pc)::value;
        v[t][p] = pair_mode_q<L, Q, p>(A, B, Q * p + qs[t]);
      });
    });
    // keep the scheduler from sinking the later slots' loads below the
    // first slot's DFT
    __builtin_amdgcn_sched_barrier(0);
    sfor<0, NS>([&](auto sc) {
      constexpr int t = decltype(sc)::value;
      const int qq = qs[t];
      dft<true, P>(v[t]);
      if (t * W + W <= Q * R || t * W + lane < Q * R) {
        C* s = S + (rs[t] * 3 + tr) * SP;
        s[qq] = v[t][0];
        sfor<1, P>([&](auto kc) {
          constexpr int k1 = decltype(kc)::value;
          s[k1 * (Q + 1) + qq] = mul_tw<true>(v[t][k1], twg[qq * k1]);
        });
      }
    });
  }
  __syncthreads();

  // phase 2: task (k1, r) of transform tr per thread; inverse Q-point DFT ->
  // physical values at x = k1 + P k2, kept in registers; the imaginary parts
  // of transforms 0 and 1 (w, then v) are written back in place for the
  // products of the next transform.  In FP32 the twiddles of phase 3 are
  // loaded first, so that they arrive during the DFT (in FP64 they would
  // push the registers past the 256 that two blocks per CU allow).
  constexpr bool kTw3Early = sizeof(T) == 4;
  const bool act = lane < P * R;
  const int k1 = lane % P, r = act ? lane / P : 0;
  C* s = S + (r * 3 + tr) * SP + k1 * (Q + 1);
  C u[Q], tw3[Q];
  if (act) {
    if constexpr (kTw3Early) sfor<1, Q>([&](auto c) { tw3[decltype(c)::value] = twg[k1 * decltype(c)::value]; });
    sfor<0, Q>([&](auto c) { u[decltype(c)::value] = s[decltype(c)::value]; });
    dft<true, Q>(u);
    if (tr < 2)
      sfor<0, Q>([&](auto c) { reinterpret_cast<T*>(s + decltype(c)::value)[1] = u[decltype(c)::value].y; });
  }
  __syncthreads();

  // products in physical space (ops_in_planes, FOU3D.f90:1529-1617); on CFL
  // substeps the maxima of |u|, |w| and |v| dtheta/dy (dtc_calc,
  // FOU3D.f90:516-544)
  T ma = 0, mb = 0;
  if (act) {
    const C* sp = S + (r * 3 + (tr > 0 ? tr - 1 : 0)) * SP + k1 * (Q + 1);  // transform tr-1
    if (a.cfl) {
      sfor<0, Q>([&](auto c) {
        const C x = u[decltype(c)::value];
        ma = fmax(ma, fabs(x.x));
        mb = fmax(mb, fabs(x.y));
      });
    }
    if (tr == 0) {
      sfor<0, Q>([&](auto c) {
        const C x = u[decltype(c)::value];
        u[decltype(c)::value] = mk<C>(x.x * x.x, x.x * x.y);  // uu*uu, uu*ww
      });
    } else if (tr == 1) {
      sfor<0, Q>([&](auto c) {
        const C x = u[decltype(c)::value];
        const T ww = sp[decltype(c)::value].y;
        u[decltype(c)::value] = mk<C>(x.x * x.x, ww * ww);  // vc*vc, ww*ww
      });
    } else {
      sfor<0, Q>([&](auto c) {
        const C x = u[decltype(c)::value];
        const T vf = sp[decltype(c)::value].y;
        u[decltype(c)::value] = mk<C>(x.x * vf, x.y * vf);  // uf*vf, wf*vf
      });
    }
  }
  // W is a multiple of 64, so a.cfl and tr are uniform across each wavefront
  // and every lane takes part in its reduction
  if (a.cfl && tr < 2) {
    double da = ma, db = mb;
    if (!act || z0 + r >= ngz) da = db = 0.0;
    if (tr == 1) db *= vint ? a.dthdyv[q] : 0.0;
    if (tr == 0 && a.cfu && !a.cfu[q]) da = db = 0.0;
    if (tr == 1 && a.cfv && !(hasv && a.cfv[q])) db = 0.0;
    da = wave_max(da);
    db = wave_max(db);
    if ((lane & 63) == 0) {
      if (tr == 0) {
        atomic_max_pos(&a.sc->m1, da);
        atomic_max_pos(&a.sc->m3, db);
      } else {
        atomic_max_pos(&a.sc->m2, db);
      }
    }
  }
  __syncthreads();

  // phase 3: forward Q-point DFT over k2 -> m2, twiddle w_L^{k1 m2}
  C* sf = S + (r * 3 + tr) * SP;
  if (act) {
    dft<false, Q>(u);
    sf[k1] = u[0];
    sfor<1, Q>([&](auto c) {
      constexpr int m2c = decltype(c)::value;
      sf[m2c * (P + 1) + k1] = mul_tw<false>(u[m2c], kTw3Early ? tw3[m2c] : twg[k1 * m2c]);
    });
  }
  __syncthreads();

  // phase 4: tasks (m2, r) of transform tr; forward P-point DFTs over k1 ->
  // F[m2 + Q m1]
  for (int j = lane; j < Q * R; j += W) {
    const int m2 = j % Q, r4 = j / Q;
    C* s4 = S + (r4 * 3 + tr) * SP + m2 * (P + 1);
    C v[P];
    sfor<0, P>([&](auto c) { v[decltype(c)::value] = s4[decltype(c)::value]; });
    dft<false, P>(v);
    // only the kept modes m < Mx and their mirrors L - m are read below:
    // outputs m1 with m2 + Q m1 in [Mx, L - Mx] for every m2 are dropped,
    // and the compiler drops the arithmetic that only they need
    sfor<0, P>([&](auto c) {
      constexpr int m1 = decltype(c)::value;
      if constexpr (!(Q * m1 >= Mx && Q * m1 + Q - 1 <= L - Mx)) s4[m1] = v[m1];
    });
  }
  __syncthreads();

  // split the spectra of a + i b into the modes of a and b, scaled by 1/L
  if (tr < 2 ? !uint : !hasv) return;
  const T hs = (T)(0.5 / L);
  C* oa = (C*)(tr == 0 ? a.zo[0] : (tr == 1 ? a.zo[3] : a.zo[4]));  // uu, vv, uv
  C* ob = (C*)(tr == 0 ? a.zo[1] : (tr == 1 ? a.zo[2] : a.zo[5]));  // uw, ww, vw
  // lane -> mode m of a chunk of W modes, all R rows per lane; F[m] sits at
  // slot (m % Q)(P+1) + m / Q, and its mirror L - m at (Q - m % Q)(P+1) +
  // P-1 - m / Q (at P - m / Q when m % Q = 0; m = 0 pairs with itself)
  sfor<0, (Mx + W - 1) / W>([&](auto cc) {
    constexpr int c0 = decltype(cc)::value * W;
    const int m = c0 + lane;
    if (c0 + W <= Mx || m < Mx) {
      const int m2 = m % Q, m1 = m / Q;
      const int fa = m2 * (P + 1) + m1;
      const int gb = m == 0 ? 0 : (m2 == 0 ? P - m1 : (Q - m2) * (P + 1) + P - 1 - m1);
      sfor<0, R>([&](auto rc) {
        constexpr int r5 = decltype(rc)::value;
        const int z = z0 + r5;
        if (z < ngz) {
          const C* F = S + (r5 * 3 + tr) * SP;
          const C f = F[fa], g = F[gb];
          const size_t ro = ((size_t)q * ngz + z) * a.zs + m;
          oa[ro] = mk<C>(hs * (f.x + g.x), hs * (f.y - g.y));
          ob[ro] = mk<C>(hs * (f.y + g.y), -hs * (f.x - g.x));
        }
      });
    }
  });
}

template <typename T, int L, int P, int Q, int R>
__global__ void __launch_bounds__(3 * xp_width(P, Q, R)) k_xpass4(XNLArgs a, const cplx_t<T>* __restrict__ twg) {
  xpass4_body<T, L, P, Q, R>(a, twg, blockIdx.x, blockIdx.y);
}

template <typename T, int L, int P, int Q, int R>
__global__ void __launch_bounds__(3 * xp_width(P, Q, R)) k_xpass4_b(const XNLArgs* __restrict__ as, BMap bm,
                                                                    const cplx_t<T>* __restrict__ twg) {
  int bx = 0, by = 0;
  const int m = bmap_member(bm, blockIdx.x, bx, by);
  if (m < 0) return;
  // read from the table in place: a private copy of the pointer arrays
  // would be indexed at run time and land in scratch memory
  xpass4_body<T, L, P, Q, R>(as[m], twg, bx, by);
}

// ---------------------------------------------------------------------------
// Four-step IB transforms (FP64): the x-pass's phases without the products.
// ---------------------------------------------------------------------------
// x-inverse of row pairs (r + i u): grid (ceil(Ngz/R), njobs), W lanes per
// block; lane task (qq, r) in phase 1, (k1, r) in phase 2, which stores the
// physical row: real part to pa, imaginary part to pb.  Rows of Mx modes.
template <int L, int P, int Q, int R>
__global__ void __launch_bounds__(xp_width(P, Q, R)) k_xinv4(const XJob* __restrict__ jobs,
                                                            const double2* __restrict__ twg, int ngz,
                                                            const double2* __restrict__ zrow) {
  using C = double2;
  constexpr int Mx = Dealias<L>::Mx, W = xp_width(P, Q, R), NS = (Q * R + W - 1) / W;
  __shared__ C S[R * P * (Q + 1)];
  const XJob J = jobs[blockIdx.y];
  const int lane = threadIdx.x, z0 = blockIdx.x * R;
  sfor<0, NS>([&](auto sc) {
    constexpr int t = decltype(sc)::value;
    const int jt = t * W + lane;
    const int j = (t * W + W <= Q * R || jt < Q * R) ? jt : Q * R - 1;
    const int qq = j % Q, r = j / Q, z = z0 + r;
    const C *A = zrow, *B = zrow;
    if (z < ngz) {
      const size_t ro = ((size_t)J.plane * ngz + z) * Mx;
      A = J.za + ro;
      B = J.zb + ro;
    }
    C v[P];
    sfor<0, P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      v[p] = pair_mode_q<L, Q, p>(A, B, Q * p + qq);
    });
    dft<true, P>(v);
    if (t * W + W <= Q * R || jt < Q * R) {
      C* s = S + r * P * (Q + 1);
      s[qq] = v[0];
      sfor<1, P>([&](auto kc) {
        constexpr int k1 = decltype(kc)::value;
        s[k1 * (Q + 1) + qq] = mul_tw<true>(v[k1], twg[qq * k1]);
      });
    }
  });
  __syncthreads();
  if (lane >= P * R) return;
  const int k1 = lane % P, r = lane / P, z = z0 + r;
  const C* s = S + r * P * (Q + 1) + k1 * (Q + 1);
  C u[Q];
  sfor<0, Q>([&](auto c) { u[decltype(c)::value] = s[decltype(c)::value]; });
  dft<true, Q>(u);
  if (z >= ngz) return;
  const size_t o = ((size_t)J.plane * ngz + z) * L + k1;
  sfor<0, Q>([&](auto c) {
    constexpr int k2 = decltype(c)::value;
    J.pa[o + P * k2] = u[k2].x;
    J.pb[o + P * k2] = u[k2].y;
  });
}

// x-forward of physical rows z, z+1 as one complex row (a + i b): grid
// (ceil(Ngz/2/R), njobs); lane (k1, r) loads and transforms over k2, then
// tasks (m2, r), then the split into the Mx modes of rows z and z+1.
template <int L, int P, int Q, int R>
__global__ void __launch_bounds__(xp_width(P, Q, R)) k_xfwd4(const XJob* __restrict__ jobs,
                                                            const double2* __restrict__ twg, int ngz) {
  using C = double2;
  constexpr int Mx = Dealias<L>::Mx, W = xp_width(P, Q, R);
  constexpr int SP = cmax(P * (Q + 1), Q * (P + 1));
  __shared__ C S[R * SP];
  const XJob J = jobs[blockIdx.y];
  const int lane = threadIdx.x, p0 = blockIdx.x * R, npair = ngz / 2;
  if (lane < P * R) {
    const int k1 = lane % P, r = lane / P, zp = p0 + r;
    C u[Q];
    if (zp < npair) {
      const double* a = J.pa + ((size_t)J.plane * ngz + 2 * zp) * L + k1;
      sfor<0, Q>([&](auto c) {
        constexpr int k2 = decltype(c)::value;
        u[k2] = make_double2(a[P * k2], a[L + P * k2]);
      });
    } else {
      sfor<0, Q>([&](auto c) { u[decltype(c)::value] = make_double2(0.0, 0.0); });
    }
    dft<false, Q>(u);
    C* sf = S + r * SP;
    sf[k1] = u[0];
    sfor<1, Q>([&](auto c) {
      constexpr int m2 = decltype(c)::value;
      sf[m2 * (P + 1) + k1] = mul_tw<false>(u[m2], twg[k1 * m2]);
    });
  }
  __syncthreads();
  for (int j = lane; j < Q * R; j += W) {
    const int m2 = j % Q, r4 = j / Q;
    C* s4 = S + r4 * SP + m2 * (P + 1);
    C v[P];
    sfor<0, P>([&](auto c) { v[decltype(c)::value] = s4[decltype(c)::value]; });
    dft<false, P>(v);
    sfor<0, P>([&](auto c) {
      constexpr int m1 = decltype(c)::value;
      if constexpr (!(Q * m1 >= Mx && Q * m1 + Q - 1 <= L - Mx)) s4[m1] = v[m1];
    });
  }
  __syncthreads();
  const double hs = 0.5 / L;
  for (int e = lane; e < Mx * R; e += W) {
    const int m = e % Mx, r5 = e / Mx, zp = p0 + r5;
    if (zp >= npair) continue;
    const int m2 = m % Q, m1 = m / Q;
    const int fa = m2 * (P + 1) + m1;
    const int gb = m == 0 ? 0 : (m2 == 0 ? P - m1 : (Q - m2) * (P + 1) + P - 1 - m1);
    const C* F = S + r5 * SP;
    const C f = F[fa], g = F[gb];
    C* o = J.zo + ((size_t)J.plane * ngz + 2 * zp) * Mx + m;
    o[0] = make_double2(hs * (f.x + g.x), hs * (f.y - g.y));
    o[Mx] = make_double2(hs * (f.y + g.y), -hs * (f.x - g.x));
  }
}

// plain z-forward of a Z plane (rows of Mx) to the N2 dealiased modes,
// scaled by 1/L: grid (ceil(Mx/TI), njobs), the z-forward's four-step
template <int L, int P, int Q, int TI>
__global__ void __launch_bounds__(TI * cmax(P, Q)) k_zfwdp4(const ZFwdJob* __restrict__ jobs,
                                                           const double2* __restrict__ twg, int Mx) {
  using C = double2;
  constexpr int N2 = Dealias<L>::N, H = Dealias<L>::H;
  __shared__ C X[P * Q * TI];
  const ZFwdJob J = jobs[blockIdx.y];
  const int tid = threadIdx.x, i0 = blockIdx.x * TI;
  const C* src = J.src + (size_t)J.ps * L * Mx;
  if (tid < Q * TI) {
    const int t = tid % TI, q = tid / TI, i = i0 + t;
    C v[P];
    sfor<0, P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      v[p] = i < Mx ? src[(size_t)(Q * p + q) * Mx + i] : make_double2(0.0, 0.0);
    });
    dft<false, P>(v);
    X[q * TI + t] = v[0];
    sfor<1, P>([&](auto kc) {
      constexpr int kk = decltype(kc)::value;
      X[(kk * Q + q) * TI + t] = mul_tw<false>(v[kk], twg[q * kk]);
    });
  }
  __syncthreads();
  const int t2 = tid % TI, k1 = tid / TI, i2 = i0 + t2;
  if (tid < P * TI && i2 < Mx) {
    C u[Q];
    sfor<0, Q>([&](auto qc) { u[decltype(qc)::value] = X[(k1 * Q + decltype(qc)::value) * TI + t2]; });
    dft<false, Q>(u);
    const double sc = 1.0 / L;
    const size_t ncol = (size_t)N2 * Mx;
    sfor<0, Q>([&](auto c) {
      constexpr int k2 = decltype(c)::value;
      if constexpr (!(P * k2 >= H && P * k2 + P <= L - H)) {
        const int sl = k1 + P * k2;
        if (sl < H || sl >= L - H) {
          const int k = sl < H ? sl : sl - L + N2;
          J.dst[(size_t)J.pd * ncol + (size_t)k * Mx + i2] = make_double2(u[k2].x * sc, u[k2].y * sc);
        }
      }
    });
  }
}

// raise the dynamic-LDS limit of a kernel once per size
void set_lds(const void* kern, size_t bytes) {
  static std::vector<std::pair<const void*, size_t>> done;
  for (auto& d : done)
    if (d.first == kern && d.second >= bytes) return;
  (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  done.push_back({kern, bytes});
}

// ---------------------------------------------------------------------------
// launch helpers per compile-time length
// ---------------------------------------------------------------------------
template <int LC>
void zinv_t(Solver& s, const ZInvJob* jobs, int njobs) {
  constexpr int TI = LC > 0 ? ti_for(LC) : kTIFallback;
  const size_t lds = lds_bytes<LC>(s.pz.L, TI);
  set_lds((const void*)k_zinv<LC, TI>, lds);
  dim3 grid((s.Mx + TI - 1) / TI, njobs);
  hipLaunchKernelGGL((k_zinv<LC, TI>), grid, dim3(kThreads), lds, s.stream, jobs, s.pz, s.d_twz, s.N2, s.Mx);
}

template <int LC>
void xpass_nl_t(Solver& s, const XNLArgs& a) {
  const size_t lds = lds_bytes<LC>(s.px.L, 3);
  set_lds((const void*)k_xpass_nl<LC>, lds);
  hipLaunchKernelGGL((k_xpass_nl<LC>), dim3(s.ngz, win_count(a.win)), dim3(kThreads), lds, s.stream, a, s.px,
                     s.d_twx);
}

template <int LC>
void zfwd_nl_t(Solver& s, const ZNLArgs& a) {
  constexpr int TI = LC > 0 ? ti_for(LC) : kTIFallback;
  static_assert(TI <= 16, "tile");
  const size_t lds = lds_bytes<LC>(s.pz.L, TI);
  set_lds((const void*)k_zfwd_nl<LC, TI>, lds);
  dim3 grid((s.Mx + TI - 1) / TI, win_count(a.winu) + win_count(a.winv));
  hipLaunchKernelGGL((k_zfwd_nl<LC, TI>), grid, dim3(kThreads), lds, s.stream, a, s.pz, s.d_twz);
}

template <int LC>
void xinv2_t(Solver& s, const XJob* jobs, int njobs) {
  const size_t lds = lds_bytes<LC>(s.px.L, 1);
  set_lds((const void*)k_xpass_inv2<LC>, lds);
  hipLaunchKernelGGL((k_xpass_inv2<LC>), dim3(s.ngz, njobs), dim3(kThreads), lds, s.stream, jobs, s.px, s.d_twx,
                     s.Mx);
}

template <int LC>
void xfwd2_t(Solver& s, const XJob* jobs, int njobs) {
  const size_t lds = lds_bytes<LC>(s.px.L, 1);
  set_lds((const void*)k_xpass_fwd2<LC>, lds);
  hipLaunchKernelGGL((k_xpass_fwd2<LC>), dim3(s.ngz / 2, njobs), dim3(kThreads), lds, s.stream, jobs, s.px,
                     s.d_twx, s.Mx);
}

template <int LC>
void zfwdp_t(Solver& s, const ZFwdJob* jobs, int njobs) {
  constexpr int TI = LC > 0 ? ti_for(LC) : kTIFallback;
  const size_t lds = lds_bytes<LC>(s.pz.L, TI);
  set_lds((const void*)k_zfwd_plain<LC, TI>, lds);
  dim3 grid((s.Mx + TI - 1) / TI, njobs);
  hipLaunchKernelGGL((k_zfwd_plain<LC, TI>), grid, dim3(kThreads), lds, s.stream, jobs, s.pz, s.d_twz, s.N2,
                     s.Mx);
}

// four-step specializations: (L, P, Q, TI) of the z kernels, (L, P, Q, R)
// of the x-pass.  The z-inverse takes Q | N2/2 and the z-forward P | N2/2
// where the length allows, so whole groups of the empty band drop out.
#ifndef CFD_ZINV4
#define CFD_ZINV4(X) X(48, 6, 8, 16) X(64, 8, 8, 16) X(72, 6, 12, 16) X(96, 6, 16, 16) X(128, 8, 16, 16) \
  X(192, 12, 16, 16) X(384, 24, 16, 8)
#endif
#ifndef CFD_ZFWD4_F64
#define CFD_ZFWD4_F64(X) X(48, 8, 6, 16) X(64, 8, 8, 16) X(72, 8, 9, 16) X(96, 8, 12, 16) X(128, 8, 16, 16) \
  X(192, 16, 12, 16) X(384, 24, 16, 8)
#endif
#ifndef CFD_ZFWD4_F32
#define CFD_ZFWD4_F32(X) CFD_ZFWD4_F64(X)
#endif
#ifndef CFD_XPASS4_F64
#define CFD_XPASS4_F64(X) X(48, 8, 6, 8) X(64, 8, 8, 8) X(72, 12, 6, 4) X(96, 8, 12, 8) X(128, 8, 16, 8) \
  X(144, 12, 12, 4) X(192, 16, 12, 4) X(256, 16, 16, 4) X(384, 16, 24, 4) X(768, 32, 24, 2)
#endif
// FP32 at 192 points: (8, 24, 8) fills every lane in both phases (three 8-point
// slots, one 24-point task per lane) in 41 KB of LDS per block; in FP64 that
// block would need 83 KB (one block per CU), so FP64 keeps (16, 12, 4)
#ifndef CFD_XPASS4_F32
#define CFD_XPASS4_F32(X) X(48, 8, 6, 8) X(64, 8, 8, 8) X(72, 12, 6, 4) X(96, 8, 12, 8) X(128, 8, 16, 8) \
  X(144, 12, 12, 4) X(192, 8, 24, 8) X(256, 16, 16, 4) X(384, 16, 24, 4) X(768, 32, 24, 2)
#endif

bool four_step_z(const Solver& s) {
  return (s.mixed || !s.fft_stockham) && s.N2 == 2 * (s.ngz / 3) && s.pz.L == s.ngz;
}

template <typename T>
bool zinv4_t(Solver& s, const ZInvJob* jobs, int njobs, const cplx_t<T>* tw, int zs, bool xcd) {
  switch (s.pz.L) {
#define C(L, P, Q, TI)                                                                                 \
  case L: {                                                                                            \
    const int nt = (s.Mx + TI - 1) / TI, jper = xcd ? (njobs + 7) / 8 : 0;                             \
    const dim3 grid = xcd ? dim3(8 * jper * nt) : dim3(nt, njobs);                                     \
    hipLaunchKernelGGL((k_zinv4<T, L, P, Q, TI>), grid, dim3(TI * cmax(P, Q)), 0, s.stream, jobs, tw, \
                       s.Mx, zs, njobs, nt, jper);                                                     \
    return true;                                                                                       \
  }
    CFD_ZINV4(C)
#undef C
    default: return false;
  }
}

bool try_zinv4(Solver& s, const ZInvJob* jobs, int njobs, bool nl) {
  if (!four_step_z(s)) return false;
  // nonlinear planes: precision of the path, padded rows, XCD-aware order;
  // IB planes: FP64, rows of Mx
  if (nl && s.mixed) return zinv4_t<float>(s, jobs, njobs, s.d_twzf, s.zs, true);
  return zinv4_t<double>(s, jobs, njobs, s.d_twz, nl ? s.zs : s.Mx, nl);
}

#define CFD_ZF_CASE(T, L, P, Q, TI)                                                                \
  case L:                                                                                       \
    hipLaunchKernelGGL((k_zfwd4<T, L, P, Q, TI>),                                              \
                       dim3((s.Mx + TI - 1) / TI, win_count(a.winu) + win_count(a.winv)),      \
                       dim3(TI * cmax(P, Q)), 0, s.stream, a, tw);                             \
    return true;

bool zfwd4_t(Solver& s, const ZNLArgs& a, const double2* tw) {
  switch (s.pz.L) {
#define C(L, P, Q, TI) CFD_ZF_CASE(double, L, P, Q, TI)
    CFD_ZFWD4_F64(C)
#undef C
    default: return false;
  }
}

bool zfwd4_t(Solver& s, const ZNLArgs& a, const float2* tw) {
  switch (s.pz.L) {
#define C(L, P, Q, TI) CFD_ZF_CASE(float, L, P, Q, TI)
    CFD_ZFWD4_F32(C)
#undef C
    default: return false;
  }
}

bool try_zfwd4(Solver& s, const ZNLArgs& a) {
  if (!four_step_z(s)) return false;
  return s.mixed ? zfwd4_t(s, a, s.d_twzf) : zfwd4_t(s, a, s.d_twz);
}

#define CFD_XP_CASE(T, L, P, Q, R)                                                                  \
  case L:                                                                                         \
    hipLaunchKernelGGL((k_xpass4<T, L, P, Q, R>), dim3((s.ngz + R - 1) / R, win_count(a.win)), \
                       dim3(3 * xp_width(P, Q, R)), 0, s.stream, a, tw);                            \
    return true;

bool xpass4_t(Solver& s, const XNLArgs& a, const double2* tw) {
  switch (s.px.L) {
#define C(L, P, Q, R) CFD_XP_CASE(double, L, P, Q, R)
    CFD_XPASS4_F64(C)
#undef C
    default: return false;
  }
}

bool xpass4_t(Solver& s, const XNLArgs& a, const float2* tw) {
  switch (s.px.L) {
#define C(L, P, Q, R) CFD_XP_CASE(float, L, P, Q, R)
    CFD_XPASS4_F32(C)
#undef C
    default: return false;
  }
}

bool try_xpass4(Solver& s, const XNLArgs& a) {
  if ((s.fft_stockham && !s.mixed) || s.px.L != s.ngx || s.N1 != 2 * (s.ngx / 3) || s.Mx != s.N1 / 2 + 1)
    return false;
  return s.mixed ? xpass4_t(s, a, s.d_twxf) : xpass4_t(s, a, s.d_twx);
}

// transform lengths with compile-time plans; other lengths use the runtime plan
#define CFD_FFT_SIZES(X) X(48) X(64) X(72) X(96) X(128) X(192) X(256) X(384) X(768)

}  // namespace

bool fft4_z_supported(int L, bool f32) {
  switch (L) {
#define C(L, P, Q, TI) case L:
    CFD_ZINV4(C)
#undef C
    break;
    default: return false;
  }
  if (f32) {
    switch (L) {
#define C(L, P, Q, TI) case L:
      CFD_ZFWD4_F32(C)
#undef C
      return true;
      default: return false;
    }
  }
  switch (L) {
#define C(L, P, Q, TI) case L:
    CFD_ZFWD4_F64(C)
#undef C
    return true;
    default: return false;
  }
}

bool fft4_x_supported(int L, bool f32) {
  if (f32) {
    switch (L) {
#define C(L, P, Q, R) case L:
      CFD_XPASS4_F32(C)
#undef C
      return true;
      default: return false;
    }
  }
  switch (L) {
#define C(L, P, Q, R) case L:
    CFD_XPASS4_F64(C)
#undef C
    return true;
    default: return false;
  }
}

void launch_zinv(Solver& s, const ZInvJob* jobs, int njobs, bool nl) {
  if (njobs <= 0) return;
  if (try_zinv4(s, jobs, njobs, nl)) return;
  switch (s.pz.L) {
#define C(n) case n: zinv_t<n>(s, jobs, njobs); break;
    CFD_FFT_SIZES(C)
#undef C
    default: zinv_t<0>(s, jobs, njobs);
  }
}

namespace {
void xpass_args(Solver& s, int cfl, XNLArgs& a) {
  for (int c = 0; c < 6; ++c) {
    a.zi[c] = s.Z[c];
    a.zo[c] = s.Z[6 + c];
  }
  a.dthdyv = s.d_dthdyv;
  a.sc = s.d_sc;
  a.nyu = s.nyu;
  a.nyv = s.nyv;
  a.Mx = s.Mx;
  a.ngz = s.ngz;
  a.cfl = cfl;
  a.zs = s.zs;
  a.top = s.compat_np > 0;
  a.cfu = s.d_cfu;
  a.cfv = s.d_cfv;
  for (int w = 0; w < 4; ++w) a.win[w] = s.nl_x[w];
  a.zrow = s.d_zrow;
}

void zfwd_args(Solver& s, ZNLArgs& a) {
  for (int c = 0; c < 6; ++c) a.z[c] = s.Z[6 + c];
  a.nxz1 = s.nxz1;
  a.nxz3 = s.nxz3;
  a.fvv = s.fvv;
  a.nxz2 = s.nxz2;
  a.fuv = s.fuv;
  a.fvw = s.fvw;
  a.k1x = s.d_k1x;
  a.k1z = s.d_k1z;
  a.nyu = s.nyu;
  a.nyv = s.nyv;
  a.N2 = s.N2;
  a.Mx = s.Mx;
  a.zs = s.zs;
  a.top = s.compat_np > 0;
  for (int w = 0; w < 4; ++w) {
    a.winu[w] = s.nl_u[w];
    a.winv[w] = s.nl_v[w];
  }
}
}  // namespace

void launch_xpass_nl(Solver& s, int cfl) {
  XNLArgs a;
  xpass_args(s, cfl, a);
  if (try_xpass4(s, a)) return;
  switch (s.px.L) {
#define C(n) case n: xpass_nl_t<n>(s, a); break;
    CFD_FFT_SIZES(C)
#undef C
    default: xpass_nl_t<0>(s, a);
  }
}

void launch_zfwd_nl(Solver& s) {
  ZNLArgs a;
  zfwd_args(s, a);
  if (try_zfwd4(s, a)) return;
  switch (s.pz.L) {
#define C(n) case n: zfwd_nl_t<n>(s, a); break;
    CFD_FFT_SIZES(C)
#undef C
    default: zfwd_nl_t<0>(s, a);
  }
}

namespace {
// the four-step IB kernels serve the lengths and dealiasing of the
// four-step x-pass and z passes (FP64 tables); the Stockham engine keeps
// its own kernels
bool ib_four_step_x(const Solver& s) {
  return !s.fft_stockham && s.px.L == s.ngx && s.N1 == 2 * (s.ngx / 3) && s.Mx == s.N1 / 2 + 1;
}
bool ib_four_step_z(const Solver& s) { return !s.fft_stockham && s.N2 == 2 * (s.ngz / 3) && s.pz.L == s.ngz; }
}  // namespace

void launch_xpass_inv2(Solver& s, const XJob* jobs, int njobs) {
  if (njobs <= 0) return;
  if (ib_four_step_x(s)) {
    switch (s.px.L) {
#define C(L, P, Q, R)                                                                                          \
  case L:                                                                                                      \
    hipLaunchKernelGGL((k_xinv4<L, P, Q, R>), dim3((s.ngz + R - 1) / R, njobs), dim3(xp_width(P, Q, R)), 0,    \
                       s.stream, jobs, s.d_twx, s.ngz, (const double2*)s.d_zrow);                              \
    return;
      CFD_XPASS4_F64(C)
#undef C
      default: break;
    }
  }
  switch (s.px.L) {
#define C(n) case n: xinv2_t<n>(s, jobs, njobs); break;
    CFD_FFT_SIZES(C)
#undef C
    default: xinv2_t<0>(s, jobs, njobs);
  }
}

void launch_xpass_fwd2(Solver& s, const XJob* jobs, int njobs) {
  if (njobs <= 0) return;
  if (ib_four_step_x(s) && s.ngz % 2 == 0) {
    switch (s.px.L) {
#define C(L, P, Q, R)                                                                                              \
  case L:                                                                                                          \
    hipLaunchKernelGGL((k_xfwd4<L, P, Q, R>), dim3((s.ngz / 2 + R - 1) / R, njobs), dim3(xp_width(P, Q, R)), 0,    \
                       s.stream, jobs, s.d_twx, s.ngz);                                                            \
    return;
      CFD_XPASS4_F64(C)
#undef C
      default: break;
    }
  }
  switch (s.px.L) {
#define C(n) case n: xfwd2_t<n>(s, jobs, njobs); break;
    CFD_FFT_SIZES(C)
#undef C
    default: xfwd2_t<0>(s, jobs, njobs);
  }
}

void launch_zfwd_plain(Solver& s, const ZFwdJob* jobs, int njobs) {
  if (njobs <= 0) return;
  if (ib_four_step_z(s)) {
    switch (s.pz.L) {
#define C(L, P, Q, TI)                                                                                     \
  case L:                                                                                                  \
    hipLaunchKernelGGL((k_zfwdp4<L, P, Q, TI>), dim3((s.Mx + TI - 1) / TI, njobs), dim3(TI * cmax(P, Q)), \
                       0, s.stream, jobs, s.d_twz, s.Mx);                                                  \
    return;
      CFD_ZFWD4_F64(C)
#undef C
      default: break;
    }
  }
  switch (s.pz.L) {
#define C(n) case n: zfwdp_t<n>(s, jobs, njobs); break;
    CFD_FFT_SIZES(C)
#undef C
    default: zfwdp_t<0>(s, jobs, njobs);
  }
}

// ---------------------------------------------------------------------------
// batched launches of the nonlinear transforms (ensembles)
// ---------------------------------------------------------------------------
namespace {
template <class A>
void put_args(BatchPart& p, const A& a) {
  p.args.resize(sizeof(A));
  std::memcpy(p.args.data(), &a, sizeof(A));
}
int fft_key(bool f32, int L) { return 2 * L + (f32 ? 1 : 0); }
}  // namespace

bool batch_part_fft(Solver& s, int cls, int kRK, BatchPart& part, std::string& err) {
  part = BatchPart();
  if (s.mb || s.compat_np > 0) {
    err = "multi-band and last-plane replay solvers have no batched transforms";
    return false;
  }
  if (cls == BC_ZINV_NL) {
    if (!four_step_z(s)) {
      err = "the z-inverse of this solver is not a four-step kernel";
      return false;
    }
    const bool nl = true, f32 = s.mixed;
    ZInvArgs a;
    a.jobs = s.d_nljobs;
    a.njobs = s.n_nljobs;
    a.Mx = s.Mx;
    a.zs = s.zs;
    int TI = 0, thr = 0;
    switch (s.pz.L) {
#define C(L, P, Q, TIv) \
  case L:               \
    TI = TIv;           \
    thr = TIv * cmax(P, Q); \
    break;
      CFD_ZINV4(C)
#undef C
      default: err = "no four-step z-inverse of this length"; return false;
    }
    a.ntiles = (s.Mx + TI - 1) / TI;
    a.jper = nl ? (a.njobs + 7) / 8 : 0;
    part.gx = nl ? 8 * a.jper * a.ntiles : a.ntiles;
    part.gy = nl ? 1 : a.njobs;
    part.block = dim3(thr);
    part.key = fft_key(f32, s.pz.L);
    part.tw = f32 ? (const void*)s.d_twzf : (const void*)s.d_twz;
    put_args(part, a);
    return true;
  }
  if (cls == BC_XPASS) {
    XNLArgs a;
    xpass_args(s, kRK == 1, a);
    if ((s.fft_stockham && !s.mixed) || s.px.L != s.ngx || s.N1 != 2 * (s.ngx / 3) || s.Mx != s.N1 / 2 + 1) {
      err = "the x-pass of this solver is not a four-step kernel";
      return false;
    }
    int R = 0, thr = 0;
    if (s.mixed) {
      switch (s.px.L) {
#define C(L, P, Q, Rv) \
  case L:              \
    R = Rv;            \
    thr = 3 * xp_width(P, Q, Rv); \
    break;
        CFD_XPASS4_F32(C)
        default: break;
      }
    } else {
      switch (s.px.L) {
        CFD_XPASS4_F64(C)
#undef C
        default: break;
      }
    }
    if (!R) {
      err = "no four-step x-pass of this length";
      return false;
    }
    part.gx = (s.ngz + R - 1) / R;
    part.gy = win_count(a.win);
    part.block = dim3(thr);
    part.key = fft_key(s.mixed, s.px.L);
    part.tw = s.mixed ? (const void*)s.d_twxf : (const void*)s.d_twx;
    put_args(part, a);
    return true;
  }
  if (cls == BC_ZFWD) {
    if (!four_step_z(s)) {
      err = "the z-forward of this solver is not a four-step kernel";
      return false;
    }
    ZNLArgs a;
    zfwd_args(s, a);
    int TI = 0, thr = 0;
    if (s.mixed) {
      switch (s.pz.L) {
#define C(L, P, Q, TIv) \
  case L:               \
    TI = TIv;           \
    thr = TIv * cmax(P, Q); \
    break;
        CFD_ZFWD4_F32(C)
        default: break;
      }
    } else {
      switch (s.pz.L) {
        CFD_ZFWD4_F64(C)
#undef C
        default: break;
      }
    }
    if (!TI) {
      err = "no four-step z-forward of this length";
      return false;
    }
    part.gx = (s.Mx + TI - 1) / TI;
    part.gy = win_count(a.winu) + win_count(a.winv);
    part.block = dim3(thr);
    part.key = fft_key(s.mixed, s.pz.L);
    part.tw = s.mixed ? (const void*)s.d_twzf : (const void*)s.d_twz;
    put_args(part, a);
    return true;
  }
  err = "not a transform class";
  return false;
}

void batch_launch_fft(hipStream_t st, const BatchLaunch& b) {
  const bool f32 = b.key & 1;
  const int L = b.key >> 1;
  if (b.cls == BC_ZINV_NL) {
    const ZInvArgs* as = static_cast<const ZInvArgs*>(b.dargs);
    if (f32) {
      switch (L) {
#define C(Lv, P, Q, TI)                                                                                   \
  case Lv:                                                                                                \
    hipLaunchKernelGGL((k_zinv4_b<float, Lv, P, Q, TI>), b.grid, b.block, 0, st, as, b.map, (const float2*)b.tw); \
    return;
        CFD_ZINV4(C)
#undef C
        default: return;
      }
    }
    switch (L) {
#define C(Lv, P, Q, TI)                                                                                     \
  case Lv:                                                                                                  \
    hipLaunchKernelGGL((k_zinv4_b<double, Lv, P, Q, TI>), b.grid, b.block, 0, st, as, b.map, (const double2*)b.tw); \
    return;
      CFD_ZINV4(C)
#undef C
      default: return;
    }
  }
  if (b.cls == BC_XPASS) {
    const XNLArgs* as = static_cast<const XNLArgs*>(b.dargs);
    if (f32) {
      switch (L) {
#define C(Lv, P, Q, R)                                                                                   \
  case Lv:                                                                                               \
    hipLaunchKernelGGL((k_xpass4_b<float, Lv, P, Q, R>), b.grid, b.block, 0, st, as, b.map, (const float2*)b.tw); \
    return;
        CFD_XPASS4_F32(C)
#undef C
        default: return;
      }
    }
    switch (L) {
#define C(Lv, P, Q, R)                                                                                     \
  case Lv:                                                                                                 \
    hipLaunchKernelGGL((k_xpass4_b<double, Lv, P, Q, R>), b.grid, b.block, 0, st, as, b.map, (const double2*)b.tw); \
    return;
      CFD_XPASS4_F64(C)
#undef C
      default: return;
    }
  }
  if (b.cls == BC_ZFWD) {
    const ZNLArgs* as = static_cast<const ZNLArgs*>(b.dargs);
    if (f32) {
      switch (L) {
#define C(Lv, P, Q, TI)                                                                                   \
  case Lv:                                                                                                \
    hipLaunchKernelGGL((k_zfwd4_b<float, Lv, P, Q, TI>), b.grid, b.block, 0, st, as, b.map, (const float2*)b.tw); \
    return;
        CFD_ZFWD4_F32(C)
#undef C
        default: return;
      }
    }
    switch (L) {
#define C(Lv, P, Q, TI)                                                                                     \
  case Lv:                                                                                                  \
    hipLaunchKernelGGL((k_zfwd4_b<double, Lv, P, Q, TI>), b.grid, b.block, 0, st, as, b.map, (const double2*)b.tw); \
    return;
      CFD_ZFWD4_F64(C)
#undef C
      default: return;
    }
  }
}

}  // namespace cfd
