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
#include "fft.hpp"

namespace cfd {

namespace {

constexpr int kThreads = 256;

__device__ inline double dabs(double x) { return x < 0 ? -x : x; }

// slot of reference mode k (0-based) in the Ngz-point z transform
__device__ __forceinline__ int zslot(int k, int N2, int ngz) { return k < N2 / 2 ? k : k - N2 + ngz; }

// ---------------------------------------------------------------------------
// z-inverse: grid (ceil(Mx/TI), njobs)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_zinv(const ZInvJob* __restrict__ jobs, FftPlan P,
                                                   const double2* __restrict__ twg, int N2, int Mx, int TI) {
  extern __shared__ double2 lds[];
  const int L = P.L;
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + TI * L;
  const ZInvJob J = jobs[blockIdx.y];
  const int i0 = blockIdx.x * TI;
  const int ncol = N2 * Mx;
  load_twiddles(tw, twg, L);
  // zero the padding slots, then scatter the N2 dealiased modes
  for (int w = threadIdx.x; w < TI * L; w += blockDim.x) A[w] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int w = threadIdx.x; w < N2 * TI; w += blockDim.x) {
    const int k = w / TI, t = w - k * TI;
    const int i = i0 + t;
    if (i >= Mx) continue;
    const size_t off = (size_t)k * Mx + i;
    double2 v;
    if (J.mode == 0) {
      v = J.src[(size_t)J.pa * ncol + off];
    } else {
      const double2 a = J.src[(size_t)J.pa * ncol + off], b = J.src[(size_t)J.pb * ncol + off];
      v = make_double2(J.wa * a.x + J.wb * b.x, J.wa * a.y + J.wb * b.y);
      if (J.mode == 2) {
        const double2 c = J.src[(size_t)J.pc * ncol + off];
        v = make_double2(J.g * (c.x - v.x) + v.x, J.g * (c.y - v.y) + v.y);
      }
    }
    A[t * L + zslot(k, N2, L)] = v;
  }
  __syncthreads();
  double2* R = fft_lds<true>(A, B, tw, P, TI);
  for (int w = threadIdx.x; w < L * TI; w += blockDim.x) {
    const int z = w / TI, t = w - z * TI;
    const int i = i0 + t;
    if (i >= Mx) continue;
    J.dst[((size_t)J.pd * L + z) * Mx + i] = R[t * L + z];
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
__device__ __forceinline__ void split_mode(const double2* F, int m, int L, double sc, double2& a, double2& b) {
  const double2 f = F[m], g = F[m == 0 ? 0 : L - m];
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
  const double2* zi[6];
  double2* zo[6];
  const double* dthdyv;
  DevScalars* sc;
  int nyu, nyv, Mx, cfl;
};

__global__ void __launch_bounds__(kThreads) k_xpass_nl(XNLArgs a, FftPlan P, const double2* __restrict__ twg) {
  extern __shared__ double2 lds[];
  const int L = P.L, Mx = a.Mx;
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + 3 * L;
  __shared__ double red[kThreads / 64];
  const int z = blockIdx.x, q = blockIdx.y;
  const int ngz = gridDim.x;
  const bool hasv = q < a.nyv;
  const bool uint = q > 0 && q < a.nyu - 1;
  const bool vint = hasv && q > 0 && q < a.nyv - 1;
  load_twiddles(tw, twg, L);
  const size_t ru = ((size_t)q * ngz + z) * Mx;
  const double2* u1 = a.zi[0] + ru;
  const double2* u2i = a.zi[1] + ru;
  const double2* u3 = a.zi[2] + ru;
  const double2* u1i = hasv ? a.zi[3] + ru : nullptr;
  const double2* v = hasv ? a.zi[4] + ru : nullptr;
  const double2* u3i = hasv ? a.zi[5] + ru : nullptr;
  for (int m = threadIdx.x; m < L; m += blockDim.x) {
    A[m] = pair_mode(u1, u3, m, L, Mx);
    A[L + m] = pair_mode(u2i, v, m, L, Mx);
    A[2 * L + m] = pair_mode(u1i, u3i, m, L, Mx);
  }
  __syncthreads();
  double2* X = fft_lds<true>(A, B, tw, P, 3);
  double2* Y = (X == A) ? B : A;
  // products in physical space; CFL maxima (dtc_calc, FOU3D.f90:516-544)
  const double dv = vint ? a.dthdyv[q] : 0.0;
  double m1 = 0.0, m2 = 0.0, m3 = 0.0;
  for (int x = threadIdx.x; x < L; x += blockDim.x) {
    const double2 p0 = X[x], p1 = X[L + x], p2 = X[2 * L + x];
    const double uu = p0.x, ww = p0.y, vc = p1.x, vf = p1.y, uf = p2.x, wf = p2.y;
    Y[x] = make_double2(uu * uu, uu * ww);
    Y[L + x] = make_double2(vc * vc, ww * ww);
    Y[2 * L + x] = make_double2(uf * vf, wf * vf);
    m1 = fmax(m1, dabs(uu));
    m3 = fmax(m3, dabs(ww));
    m2 = fmax(m2, dabs(vf) * dv);
  }
  if (a.cfl) {
    m1 = block_max(m1, red);
    if (threadIdx.x == 0) atomic_max_pos(&a.sc->m1, m1);
    m3 = block_max(m3, red);
    if (threadIdx.x == 0) atomic_max_pos(&a.sc->m3, m3);
    m2 = block_max(m2, red);
    if (threadIdx.x == 0) atomic_max_pos(&a.sc->m2, m2);
  }
  __syncthreads();
  const double2* F = fft_lds<false>(Y, X, tw, P, 3);
  const double sc = 1.0 / L;
  for (int m = threadIdx.x; m < Mx; m += blockDim.x) {
    double2 ca, cb;
    if (uint) {
      split_mode(F, m, L, sc, ca, cb);
      a.zo[0][ru + m] = ca;  // uu
      a.zo[1][ru + m] = cb;  // uw
      split_mode(F + L, m, L, sc, ca, cb);
      a.zo[3][ru + m] = ca;  // vv
      a.zo[2][ru + m] = cb;  // ww
    }
    if (hasv) {
      split_mode(F + 2 * L, m, L, sc, ca, cb);
      a.zo[4][ru + m] = ca;  // uv
      a.zo[5][ru + m] = cb;  // vw
    }
  }
}

// ---------------------------------------------------------------------------
// z-forward of the nonlinear products with the x/z derivatives.
// grid (ceil(Mx/TI), nyu + nyv): y-block q < nyu is u-plane q, else v-plane.
// ---------------------------------------------------------------------------
struct ZNLArgs {
  const double2* z[6];  // uu uw ww vv (u planes), uv vw (v planes)
  double2 *nxz1, *nxz3, *fvv, *nxz2, *fuv, *fvw;
  const double *k1x, *k1z;
  int nyu, nyv, N2, Mx, TI;
};

constexpr int kMaxOut = 8;  // outputs (k,t) per thread: N2*TI <= kMaxOut*kThreads

__global__ void __launch_bounds__(kThreads) k_zfwd_nl(ZNLArgs a, FftPlan P, const double2* __restrict__ twg) {
  extern __shared__ double2 lds[];
  const int L = P.L, Mx = a.Mx, N2 = a.N2, TI = a.TI;
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + TI * L;
  const int i0 = blockIdx.x * TI;
  const bool isu = (int)blockIdx.y < a.nyu;
  const int q = isu ? blockIdx.y : blockIdx.y - a.nyu;
  const int ny = isu ? a.nyu : a.nyv;
  const bool inter = q > 0 && q < ny - 1;
  if (isu && !inter) return;  // ghost u planes: no product is needed there
  load_twiddles(tw, twg, L);
  const int ncol = N2 * Mx;
  const double sc = 1.0 / L;
  double2 acc1[kMaxOut], acc2[kMaxOut];
  for (int c = 0; c < kMaxOut; ++c) acc1[c] = acc2[c] = make_double2(0, 0);
  const int nprod = isu ? 4 : 2;
  for (int pr = 0; pr < nprod; ++pr) {
    const double2* src = a.z[isu ? pr : 4 + pr];
    for (int w = threadIdx.x; w < TI * L; w += blockDim.x) {
      const int zz = w / TI, t = w - zz * TI;
      const int i = i0 + t;
      A[t * L + zz] = (i < Mx) ? src[((size_t)q * L + zz) * Mx + i] : make_double2(0, 0);
    }
    __syncthreads();
    const double2* R = fft_lds<false>(A, B, tw, P, TI);
#pragma unroll
    for (int c = 0; c < kMaxOut; ++c) {
      const int w = threadIdx.x + c * blockDim.x;
      if (w >= N2 * TI) break;
      const int k = w / TI, t = w - k * TI;
      const int i = i0 + t;
      if (i >= Mx) continue;
      const int sl = zslot(k, N2, L);
      const double2 X = cscl(R[t * L + sl], sc);
      const bool zero = (sl == L - N2 / 2);  // der_x/der_z zero band (FOU3D.f90:329-333)
      const double kx = a.k1x[i], kz = a.k1z[k];
      const size_t o = (size_t)q * ncol + (size_t)k * Mx + i;
      if (isu) {
        // Nu1 = ikx uu + ikz uw ; Nu3 = ikx uw + ikz ww ; keep vv
        if (pr == 0) acc1[c] = imulk(kx, X);
        if (pr == 1) { acc1[c] = cadd(acc1[c], imulk(kz, X)); acc2[c] = imulk(kx, X); }
        if (pr == 2) acc2[c] = cadd(acc2[c], imulk(kz, X));
        if (pr == 3) {
          a.fvv[o] = X;
          a.nxz1[o] = zero ? make_double2(0, 0) : acc1[c];
          a.nxz3[o] = zero ? make_double2(0, 0) : acc2[c];
        }
      } else {
        // Nu2 = ikx uv + ikz vw ; keep uv, vw
        if (pr == 0) { acc1[c] = imulk(kx, X); a.fuv[o] = X; }
        if (pr == 1) {
          a.fvw[o] = X;
          if (inter) a.nxz2[o] = zero ? make_double2(0, 0) : cadd(acc1[c], imulk(kz, X));
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
__global__ void __launch_bounds__(kThreads) k_xpass_inv2(const XJob* __restrict__ jobs, FftPlan P,
                                                         const double2* __restrict__ twg, int Mx) {
  extern __shared__ double2 lds[];
  const int L = P.L;
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + L;
  const XJob J = jobs[blockIdx.y];
  const int z = blockIdx.x, ngz = gridDim.x;
  load_twiddles(tw, twg, L);
  const size_t r = ((size_t)J.plane * ngz + z) * Mx;
  for (int m = threadIdx.x; m < L; m += blockDim.x) A[m] = pair_mode(J.za + r, J.zb + r, m, L, Mx);
  __syncthreads();
  const double2* X = fft_lds<true>(A, B, tw, P, 1);
  const size_t o = ((size_t)J.plane * ngz + z) * L;
  for (int x = threadIdx.x; x < L; x += blockDim.x) {
    J.pa[o + x] = X[x].x;
    J.pb[o + x] = X[x].y;
  }
}

// forward: physical rows z, z+1 of one plane -> spectral rows; grid (Ngz/2, njobs)
__global__ void __launch_bounds__(kThreads) k_xpass_fwd2(const XJob* __restrict__ jobs, FftPlan P,
                                                         const double2* __restrict__ twg, int Mx) {
  extern __shared__ double2 lds[];
  const int L = P.L;
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + L;
  const XJob J = jobs[blockIdx.y];
  const int z = 2 * blockIdx.x, ngz = 2 * gridDim.x;
  load_twiddles(tw, twg, L);
  const size_t o = ((size_t)J.plane * ngz + z) * L;
  for (int x = threadIdx.x; x < L; x += blockDim.x) A[x] = make_double2(J.pa[o + x], J.pa[o + L + x]);
  __syncthreads();
  const double2* F = fft_lds<false>(A, B, tw, P, 1);
  const size_t r = ((size_t)J.plane * ngz + z) * Mx;
  const double sc = 1.0 / L;
  for (int m = threadIdx.x; m < Mx; m += blockDim.x) {
    double2 ca, cb;
    split_mode(F, m, L, sc, ca, cb);
    J.zo[r + m] = ca;
    J.zo[r + Mx + m] = cb;
  }
}

// plain z-forward: Z plane -> N2 dealiased spectral modes; grid (ceil(Mx/TI), njobs)
__global__ void __launch_bounds__(kThreads) k_zfwd_plain(const ZFwdJob* __restrict__ jobs, FftPlan P,
                                                         const double2* __restrict__ twg, int N2, int Mx, int TI) {
  extern __shared__ double2 lds[];
  const int L = P.L;
  double2* tw = lds;
  double2* A = lds + L;
  double2* B = A + TI * L;
  const ZFwdJob J = jobs[blockIdx.y];
  const int i0 = blockIdx.x * TI;
  load_twiddles(tw, twg, L);
  for (int w = threadIdx.x; w < TI * L; w += blockDim.x) {
    const int zz = w / TI, t = w - zz * TI;
    const int i = i0 + t;
    A[t * L + zz] = (i < Mx) ? J.src[((size_t)J.ps * L + zz) * Mx + i] : make_double2(0, 0);
  }
  __syncthreads();
  const double2* R = fft_lds<false>(A, B, tw, P, TI);
  const double sc = 1.0 / L;
  const int ncol = N2 * Mx;
  for (int w = threadIdx.x; w < N2 * TI; w += blockDim.x) {
    const int k = w / TI, t = w - k * TI;
    const int i = i0 + t;
    if (i >= Mx) continue;
    J.dst[(size_t)J.pd * ncol + (size_t)k * Mx + i] = cscl(R[t * L + zslot(k, N2, L)], sc);
  }
}

int tile_for(int L) {
  int TI = 8;
  while (TI > 1 && (size_t)TI * L * 2 * sizeof(double2) > 48 * 1024) TI /= 2;
  return TI;
}

// raise the dynamic-LDS limit of a kernel once per size
void set_lds(const void* kern, size_t bytes) {
  static std::vector<std::pair<const void*, size_t>> done;
  for (auto& d : done)
    if (d.first == kern && d.second >= bytes) return;
  (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  done.push_back({kern, bytes});
}

}  // namespace

void launch_zinv(Solver& s, const ZInvJob* jobs, int njobs) {
  if (njobs <= 0) return;
  const int L = s.pz.L, TI = tile_for(L);
  const size_t lds = sizeof(double2) * (L + 2 * TI * L);
  set_lds((const void*)k_zinv, lds);
  dim3 grid((s.Mx + TI - 1) / TI, njobs);
  hipLaunchKernelGGL(k_zinv, grid, dim3(kThreads), lds, s.stream, jobs, s.pz, s.d_twz, s.N2, s.Mx, TI);
}

void launch_xpass_nl(Solver& s, int cfl) {
  XNLArgs a;
  for (int c = 0; c < 6; ++c) {
    a.zi[c] = s.Z[c];
    a.zo[c] = s.Z[6 + c];
  }
  a.dthdyv = s.d_dthdyv;
  a.sc = s.d_sc;
  a.nyu = s.nyu;
  a.nyv = s.nyv;
  a.Mx = s.Mx;
  a.cfl = cfl;
  const int L = s.px.L;
  const size_t lds = sizeof(double2) * (L + 6 * L);
  set_lds((const void*)k_xpass_nl, lds);
  hipLaunchKernelGGL(k_xpass_nl, dim3(s.ngz, s.nyu), dim3(kThreads), lds, s.stream, a, s.px, s.d_twx);
}

void launch_zfwd_nl(Solver& s) {
  ZNLArgs a;
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
  const int L = s.pz.L;
  int TI = tile_for(L);
  while (TI > 1 && s.N2 * TI > kMaxOut * kThreads) TI /= 2;
  a.TI = TI;
  const size_t lds = sizeof(double2) * (L + 2 * TI * L);
  set_lds((const void*)k_zfwd_nl, lds);
  dim3 grid((s.Mx + TI - 1) / TI, s.nyu + s.nyv);
  hipLaunchKernelGGL(k_zfwd_nl, grid, dim3(kThreads), lds, s.stream, a, s.pz, s.d_twz);
}

void launch_xpass_inv2(Solver& s, const XJob* jobs, int njobs) {
  if (njobs <= 0) return;
  const int L = s.px.L;
  const size_t lds = sizeof(double2) * 3 * L;
  set_lds((const void*)k_xpass_inv2, lds);
  hipLaunchKernelGGL(k_xpass_inv2, dim3(s.ngz, njobs), dim3(kThreads), lds, s.stream, jobs, s.px, s.d_twx, s.Mx);
}

void launch_xpass_fwd2(Solver& s, const XJob* jobs, int njobs) {
  if (njobs <= 0) return;
  const int L = s.px.L;
  const size_t lds = sizeof(double2) * 3 * L;
  set_lds((const void*)k_xpass_fwd2, lds);
  hipLaunchKernelGGL(k_xpass_fwd2, dim3(s.ngz / 2, njobs), dim3(kThreads), lds, s.stream, jobs, s.px, s.d_twx,
                     s.Mx);
}

void launch_zfwd_plain(Solver& s, const ZFwdJob* jobs, int njobs) {
  if (njobs <= 0) return;
  const int L = s.pz.L, TI = tile_for(L);
  const size_t lds = sizeof(double2) * (L + 2 * TI * L);
  set_lds((const void*)k_zfwd_plain, lds);
  dim3 grid((s.Mx + TI - 1) / TI, njobs);
  hipLaunchKernelGGL(k_zfwd_plain, grid, dim3(kThreads), lds, s.stream, jobs, s.pz, s.d_twz, s.N2, s.Mx, TI);
}

}  // namespace cfd
