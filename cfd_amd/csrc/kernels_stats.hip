// kernels_stats.hip -- statistics and spectra of the channel state
// (stats.f90:1-253, spectra.f90:1-59, omega_x of FOU3D.f90:187-208).
//
// Spectra are per-(plane, column) accumulations straight from the spectral
// fields.  The plane statistics need the physical planes: the host streams
// the planes of each grid through the z- and x-inverse kernels of the
// solver in chunks, and the two reduction kernels here turn each chunk into
// plane sums and tile-conditional sums.  All accumulators are FP64 and add
// one sample per call in the reference's order of operations (contraction
// into fused multiply-adds is switched off where the reference adds a
// product to a running sum).
#include <algorithm>

#include "common.hpp"

namespace cfd {

namespace {

// v interpolated to a u plane, as the nonlinear z-inverse jobs carry it
// (interp_v, FOU3D.f90:283-308): mode 1 two-point interpolation, mode 2 the
// wall-ghost extrapolation
__device__ __forceinline__ double2 interp_job(const ZInvJob& J, size_t off, size_t ncol) {
  if (J.mode == 0) return J.src[(size_t)J.pa * ncol + off];
  const double2 a = J.src[(size_t)J.pa * ncol + off], b = J.src[(size_t)J.pb * ncol + off];
  double2 v = make_double2(J.wa * a.x + J.wb * b.x, J.wa * a.y + J.wb * b.y);
  if (J.mode == 2) {
    const double2 c = J.src[(size_t)J.pc * ncol + off];
    v = make_double2(J.g * (c.x - v.x) + v.x, J.g * (c.y - v.y) + v.y);
  }
  return v;
}

// buildsp / buildspUV (spectra.f90:17-59): sp(j, column) += |u|^2 over every
// plane of the grid and every column; sp(j, column) += Re u Re v + Im u Im v
// with v = u2 interpolated to the u planes.  Grid (ceil(Ncol/256), max ny).
__global__ void __launch_bounds__(256) k_spectra(StatsSpecArgs A) {
#pragma clang fp contract(off)
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (col >= A.Ncol) return;
  const size_t e = (size_t)q * A.Ncol + col;
  if (q < A.nyu) {
    const double2 a = A.u1[e], c = A.u3[e];
    const double2 b = interp_job(A.ju2[q], col, A.Ncol);
    A.spU[e] = A.spU[e] + a.x * a.x + a.y * a.y;
    A.spW[e] = A.spW[e] + c.x * c.x + c.y * c.y;
    A.spUV[e] = A.spUV[e] + a.x * b.x + a.y * b.y;
  }
  if (q < A.nyv) {
    const double2 v = A.u2[e];
    A.spV[e] = A.spV[e] + v.x * v.x + v.y * v.y;
  }
  if (q < A.nyp) {
    const double2 v = A.p[e];
    A.spP[e] = A.spP[e] + v.x * v.x + v.y * v.y;
  }
}

// omega_x = du3/dy - i kz u2 on v planes q0 .. q0 + gridDim.y - 1
// (der_z FOU3D.f90:373-413 with the kz = -N2/2 line zeroed; der_yv_h_wx
// FOU3D.f90:495-514; wx = -wx + du3dy, FOU3D.f90:200-205), written to plane
// q - q0 of the chunk buffer.
__global__ void __launch_bounds__(256) k_omega_x(StatsWxArgs A) {
#pragma clang fp contract(off)
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= A.Ncol) return;
  const int t = blockIdx.y, q = A.q0 + t;
  const int k = col / A.Mx;
  const double kz = (k == A.N2 / 2) ? 0.0 : A.k1z[k];
  const double2 a = A.u3[(size_t)(q + 1) * A.Ncol + col], b = A.u3[(size_t)q * A.Ncol + col];
  const double m = A.dthdyv[q];
  const double2 d = make_double2((a.x - b.x) * m, (a.y - b.y) * m);
  const double2 v = A.u2[(size_t)q * A.Ncol + col];
  // (0, kz) * v = (-kz v.y, kz v.x); wx = -(that) + d
  A.wx[(size_t)t * A.Ncol + col] = make_double2(kz * v.y + d.x, -(kz * v.x) + d.y);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// plane sums (stats.f90:16-88): block (plane t of the chunk, task).  The
// task's first moment is sum a (b null) or sum a*b; with b null and m2
// set, also sum a^2.  The sums run over i < nx, k < nz of rows of ngx.
__global__ void __launch_bounds__(256) k_stat_sums(const StatSumTask* __restrict__ tasks, int ngx, size_t ppl,
                                                   int j0) {
  const StatSumTask T = tasks[blockIdx.y];
  const int t = blockIdx.x;
  const double* a = T.a + (size_t)t * ppl;
  const double* b = T.b ? T.b + (size_t)t * ppl : nullptr;
  const int n = T.nx * T.nz;
  double s1 = 0.0, s2 = 0.0;
  for (int w = threadIdx.x; w < n; w += blockDim.x) {
    const int k = w / T.nx, i = w - k * T.nx;
    const double x = a[(size_t)k * ngx + i];
    if (b) {
      s1 += x * b[(size_t)k * ngx + i];
    } else {
      s1 += x;
      s2 += x * x;
    }
  }
  __shared__ double red[2][4];
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = s1;
    red[1][wv] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r1 = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double r2 = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    T.m[j0 + t] += r1;
    if (T.m2) T.m2[j0 + t] += r2;
  }
}

// conddown1/2, condup1/2 (stats.f90:110-253): thread = tile cell (i, k),
// block y = plane t of the chunk, block z = task.  A plane of the bottom
// band adds to the bottom sums at j1 = j - jref_bot, one of the top band to
// the top sums at j1 = jref_top - j; tiles are visited iz outer, ix inner,
// each added to the running sum as the reference does.
__global__ void __launch_bounds__(256) k_stat_cond(const StatCondTask* __restrict__ tasks, int ngx, size_t ppl,
                                                   int dnx, int dnz, int ntx, int ntz, int jfirst) {
#pragma clang fp contract(off)
  const StatCondTask T = tasks[blockIdx.z];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dnx * dnz) return;
  const int t = blockIdx.y, j = jfirst + t;
  int j1;
  bool top;
  if (j >= T.bot0 && j <= T.bot1) {
    j1 = j - T.jref_bot;
    top = false;
  } else if (j >= T.top0 && j <= T.top1) {
    j1 = T.jref_top - j;
    top = true;
  } else {
    return;
  }
  const int k = c / dnx, i = c - k * dnx;
  const size_t o = (size_t)j1 * dnx * dnz + c;
  const double* a = T.a + (size_t)t * ppl;
  if (T.b) {
    const double* b = T.b + (size_t)t * ppl;
    double* m = top ? T.mt : T.mb;
    double acc = m[o];
    for (int iz = 0; iz < ntz; ++iz)
      for (int ix = 0; ix < ntx; ++ix) {
        const size_t e = (size_t)(k + dnz * iz) * ngx + i + dnx * ix;
        acc = acc + (top ? a[e] * T.sa * b[e] * T.sb : a[e] * b[e]);
      }
    m[o] = acc;
  } else {
    double* m = top ? T.mt : T.mb;
    double* m2 = top ? T.m2t : T.m2b;
    double acc = m[o], acc2 = m2[o];
    for (int iz = 0; iz < ntz; ++iz)
      for (int ix = 0; ix < ntx; ++ix) {
        const double x = a[(size_t)(k + dnz * iz) * ngx + i + dnx * ix];
        acc = acc + (top ? x * T.sa : x);
        acc2 = acc2 + x * x;
      }
    m[o] = acc;
    m2[o] = acc2;
  }
}

}  // namespace

void launch_spectra(Solver& s, const StatsSpecArgs& A) {
  const int ny = std::max(A.nyu, std::max(A.nyv, A.nyp));
  hipLaunchKernelGGL(k_spectra, dim3((A.Ncol + 255) / 256, ny), dim3(256), 0, s.stream, A);
}

void launch_omega_x(Solver& s, const StatsWxArgs& A, int nplanes) {
  if (nplanes <= 0) return;
  hipLaunchKernelGGL(k_omega_x, dim3((A.Ncol + 255) / 256, nplanes), dim3(256), 0, s.stream, A);
}

void launch_stat_sums(Solver& s, const StatSumTask* tasks, int ntasks, int nplanes, int j0) {
  if (nplanes <= 0 || ntasks <= 0) return;
  hipLaunchKernelGGL(k_stat_sums, dim3(nplanes, ntasks), dim3(256), 0, s.stream, tasks, s.ngx,
                     (size_t)s.ngz * s.ngx, j0);
}

void launch_stat_cond(Solver& s, const StatCondTask* tasks, int ntasks, int nplanes, int dnx, int dnz, int jfirst) {
  if (nplanes <= 0 || ntasks <= 0 || dnx * dnz <= 0) return;
  hipLaunchKernelGGL(k_stat_cond, dim3((dnx * dnz + 255) / 256, nplanes, ntasks), dim3(256), 0, s.stream, tasks,
                     s.ngx, (size_t)s.ngz * s.ngx, dnx, dnz, s.P.ntilex, s.P.ntilez, jfirst);
}

}  // namespace cfd
