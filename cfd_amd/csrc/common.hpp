// common.hpp -- shared declarations of the HIP rough-channel substep.
//
// Device data layout (one GPU owns the whole domain):
//   spectral field  f[plane][k][i]  double2, plane = j - jlo of its grid,
//                   k = 0..N2-1 in the reference column order (kz >= 0 first,
//                   then the negative modes), i = 0..Mx-1 (Mx = N1/2+1).
//                   Element (j, column) sits at plane*Ncol + column, column =
//                   k*Mx + i: the reference column index, so a y-line of one
//                   (kx,kz) mode is strided by Ncol and a wavefront of adjacent
//                   columns reads one contiguous run per y row.
//   z-transformed   Z[plane][z][i]  double2, z = 0..Ngz-1 (FFT order), the
//                   half-transformed planes between the z and x passes.
//   physical        P[plane][z][x]  double, x = 0..Ngx-1 (IB planes only; the
//                   nonlinear products never leave the chip in physical form).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cfd_amd.h"
#include "setup.hpp"

namespace cfd {

constexpr int kMaxRadix = 16;

struct FftPlan {
  int L;             // transform length
  int nr;            // number of radix passes
  int rad[kMaxRadix];
};

// z-inverse job: one plane of a source field (optionally interpolated in y)
// to one plane of a Z buffer.
//   mode 0: v = S[pa]
//   mode 1: v = wa*S[pa] + wb*S[pb]
//   mode 2: t = wa*S[pa] + wb*S[pb]; v = g*(S[pc] - t) + t
struct ZInvJob {
  const double2* src;
  void* dst;  // double2, or float2 for the nonlinear planes of the mixed path
  int pa, pb, pc, pd;
  int mode;
  int flags;  // bit 0: leave out the kz = -N2/2 mode (as four_to_phys's resize does)
  double wa, wb, g;
};

// x-pass job of the IB transforms: row block of one plane of a pair of Z
// buffers (inverse) or of a physical buffer (forward).
struct XJob {
  const double2* za;   // inverse: spectral rows of field a
  const double2* zb;   // inverse: spectral rows of field b
  double* pa;          // inverse: physical rows of field a / forward: rows
  double* pb;          // inverse: physical rows of field b
  double2* zo;         // forward: output spectral rows
  int plane;           // plane index inside the buffers
  int pad;
};

// z-forward job of the IB transforms: Z plane -> spectral plane (N2 modes).
struct ZFwdJob {
  const double2* src;
  double2* dst;
  int ps, pd;
};

// one entry of a device IB list (0-based in-plane indices)
struct IBDev {
  int o, o2, o3;       // z*Ngx + x of the point and of its two neighbours
  int plane;           // IB-plane index inside the physical buffers
  double lap, w2, w3, pad;
};

// device-resident scalars of the time stepping (declaration.f90:49-52, 119)
struct DevScalars {
  double t, dt, dti, mpgx, dgx, QxT, Qx, dtc;
  unsigned long long m1, m2, m3;  // CFL maxima as IEEE bit patterns (>= 0)
  unsigned long long amax;        // max |div| of the history record, bit pattern (>= 0)
};

enum KClass { KC_ZINV = 0, KC_XPASS, KC_ZFWD, KC_ELEM, KC_YSOLVE, KC_PRESS, KC_IB, KC_N };

struct Solver {
  cfd_params P{};
  std::string err;
  int dev = 0;
  hipStream_t stream = nullptr;
  // the flow-rate correction's u11 solve runs on `side` beside the column
  // solves (fork after the CFL step, join before the correction)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double* d_u11 = nullptr;

  // sizes
  int ngx, ngz, ny, dsty, N1, N2, Mx, Ncol;
  int nyu, nyv, nyp, jlo_u, jlo_v, jlo_p;
  int nn;

  // host copies of setup products
  YGrid yg;
  std::vector<IBPoint> ib[2][2];
  double gw[2], gwi[2], dtv;
  double aRK[3], bRK[3], gRK[3], cRK[3], dRK[3];
  std::vector<double> k1x, k2x, k1z, k2z;

  // device metrics
  double *d_dthdyu = nullptr, *d_dthdyv = nullptr, *d_d2u = nullptr, *d_d2v = nullptr;
  double *d_k1x = nullptr, *d_k2x = nullptr, *d_k1z = nullptr, *d_k2z = nullptr;
  double *d_plo = nullptr, *d_pup = nullptr, *d_pdg = nullptr;  // pressure operator rows
  double* d_yv = nullptr;                                       // flow-rate weights
  double* d_mf = nullptr;                                       // mean-flow scratch (4*(ny+1))

  // device fields
  double2* F[CFD_NFIELDS] = {};
  // precision of the nonlinear term: false = FP64 everywhere; true = the
  // mixed path, whose Z planes, transforms, products and nonlinear scratch
  // (nxz*, fluxes) are FP32 (float2) while fields and solves stay FP64
  bool mixed = false;
  // nonlinear scratch: x/z parts and fluxes (spectral), double2 or float2
  void *nxz1 = nullptr, *nxz2 = nullptr, *nxz3 = nullptr, *fuv = nullptr, *fvw = nullptr, *fvv = nullptr;
  // Z buffers: inverse inputs 0..5 (u1,u2i,u3 on u planes; u1i,v,u3i on v planes)
  // and products 6..11 (uu,uw,ww,vv on u planes; uv,vw on v planes), double2 or float2
  void* Z[12] = {};
  int zs = 0;  // Z row stride in elements: Mx rounded up to 128 B when the four-step kernels own Z
  double* cscr = nullptr;  // per-column factor scratch for the y solves (3 fields)
  double2* ysc[3] = {};    // forward-sweep scratch of the fused solves
  unsigned char* d_ibmask[3] = {};  // 1 on the IB planes of field f
  int* d_ibq[3] = {};               // plane indices of the IB planes of field f
  DevScalars* d_sc = nullptr;

  // FFT
  FftPlan pz{}, px{};
  bool fft_stockham = false;  // force the LDS Stockham kernels (CFD_AMD_FFT=stockham)
  // launch shapes of the segmented y solves (k_rhs_seg, k_press_seg): rows
  // per segment (5, 9, 13 or 21) and the cap on threads per block (64..256)
  int seg_rows = 13;
  int solve_threads = 256;
  long long nsub = 0;  // substeps launched (captured ones included)
  // modal filter of the velocity after every RK step (off by default):
  // sigma per column, applied to every plane of u1, u2, u3
  double* d_sigma = nullptr;
  bool filt = false;
  // the reference's last-plane nonlinear term for np ranks (0: per plane):
  // plane q of the x/z part takes plane map[q] (u and v grids)
  int compat_np = 0;
  int *d_umap = nullptr, *d_vmap = nullptr;
  double2 *d_twz = nullptr, *d_twx = nullptr;
  float2 *d_twzf = nullptr, *d_twxf = nullptr;  // the same tables rounded to FP32
  double2* d_zrow = nullptr;                      // Ngx zeros: the x-pass's stand-in for an absent row

  // job tables
  ZInvJob* d_nljobs = nullptr;
  int n_nljobs = 0;

  // IB
  int nibp[3] = {0, 0, 0};            // IB planes per field (u1, u2, u3)
  std::vector<int> ibplanes[2];       // global j of each IB plane, per grid
  double2 *zib_r[3] = {}, *zib_u[3] = {};
  double *pib_r[3] = {}, *pib_u[3] = {};
  ZInvJob* d_ibz = nullptr;
  int n_ibz = 0;
  XJob *d_ibxi = nullptr, *d_ibxf = nullptr;
  int n_ibxi = 0, n_ibxf = 0;
  ZFwdJob* d_ibzf = nullptr;
  int n_ibzf = 0;
  IBDev* d_ibl[2] = {};
  int* d_ibseg[2] = {};
  int nibseg[2] = {0, 0};

  // multi-band resolution (start.f90:89-111, 1420-1557): with mb the central
  // y band is transformed on the coarser ngx2 x ngz2 grid and holds only the
  // long columns (the N1c x N2c modes); the short columns live in the wall
  // bands only, each band a closed system whose boundary rows sit at the
  // band interface.  The central planes' nonlinear products come from mbc, a
  // solver of the central band's resolution over the long columns that
  // shares this solver's stream.
  bool mb = false;
  int ngx2 = 0, ngz2 = 0, N1c = 0, N2c = 0;
  int N3[4] = {0, 0, 0, 0}, N4[4] = {0, 0, 0, 0};  // band limits N(3,0:3), N(4,0:3)
  double gw12 = 0.0, gw31 = 0.0;  // gridweighting(1,2), (3,1): the interface boundary rows
  Solver* mbc = nullptr;
  bool borrowed_stream = false;      // a central solver runs on its parent's stream
  int* d_ccol = nullptr;             // per column: its column in mbc, -1 for a short column
  int* d_lcol = nullptr;             // per mbc column: the column here
  unsigned char* d_short = nullptr;  // per column: 1 = short
  unsigned char *d_cfu = nullptr, *d_cfv = nullptr;  // per u / v plane: 1 = in this solver's CFL maxima
  std::vector<unsigned char> h_short;
  // plane windows {lo, hi, s0, s1} of the nonlinear transforms: planes [lo,
  // hi) without [s0, s1) of the x-pass (nl_x) and of the z-forward's u and v
  // planes (nl_u, nl_v); all planes unless multi-band resolution splits them
  // between the two resolutions
  int nl_x[4] = {0, 0, 0, 0}, nl_u[4] = {0, 0, 0, 0}, nl_v[4] = {0, 0, 0, 0};
  std::vector<ZInvJob> h_nljobs;  // the z-inverse jobs of the nonlinear term (host copy)
  int gather_q[2] = {0, 0};       // multi-band: planes [q0, q1) whose long columns mbc takes

  // profiling
  bool prof = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double prof_ms[KC_N] = {};
  int prof_cnt[KC_N] = {};
};

// launchers (kernels_fft.hip)
// nl: the nonlinear planes (precision of the path, rows of zs, XCD-aware
// order); otherwise FP64 IB planes with rows of Mx
void launch_zinv(Solver& s, const ZInvJob* jobs, int njobs, bool nl = false);
// transform lengths served by the four-step kernels in FP32 (f32) or FP64;
// the mixed path needs them
bool fft4_z_supported(int L, bool f32);
bool fft4_x_supported(int L, bool f32);
void launch_xpass_nl(Solver& s, int cfl);
void launch_zfwd_nl(Solver& s);
void launch_xpass_inv2(Solver& s, const XJob* jobs, int njobs);
void launch_xpass_fwd2(Solver& s, const XJob* jobs, int njobs);
void launch_zfwd_plain(Solver& s, const ZFwdJob* jobs, int njobs);
// launchers (kernels_spec.hip)
void launch_rhs0(Solver& s, int kRK);
void launch_nlcombine(Solver& s, int kRK);
void launch_reset_max(Solver& s);
void launch_cfl(Solver& s);
void launch_predictor(Solver& s, int field);
void launch_ib_apply(Solver& s, int f0, int f1);  // IB forcing of fields [f0, f1) in one launch
void launch_ysolve(Solver& s, int field, int kRK);
void launch_meanflow(Solver& s, int kRK);
void launch_solve_p(Solver& s, int kRK);
void launch_vcorr(Solver& s, int kRK);
void launch_divergence(Solver& s);
void launch_flowrate(Solver& s);
void launch_mode_filter(Solver& s);  // u_i *= sigma(column) on every plane
void launch_plane_map(Solver& s);   // last-plane replay of nxz1, nxz3 (u) and nxz2 (v)
void launch_absmax(Solver& s, const double2* a, size_t n);  // max |a| -> d_sc->amax
bool seg_rows_supported(int rows);
// statistics (kernels_stats.hip)
struct StatsSpecArgs {
  const double2 *u1, *u2, *u3, *p;
  const ZInvJob* ju2;  // per u plane: u2 interpolated to it (the nonlinear z-inverse job)
  double *spU, *spV, *spW, *spUV, *spP;  // sp[q * Ncol + col]
  int Ncol, nyu, nyv, nyp;
};
struct StatsWxArgs {
  const double2 *u2, *u3;
  const double *dthdyv, *k1z;
  double2* wx;  // chunk buffer, plane q - q0
  int q0, Ncol, Mx, N2;
};
struct StatSumTask {
  const double* a;  // physical chunk planes (rows of ngx)
  const double* b;  // second factor or null
  double *m, *m2;   // profile accumulators (m2 may be null)
  int nx, nz;       // summed extent
};
struct StatCondTask {
  const double *a, *b;
  double *mb, *m2b, *mt, *m2t;  // bottom / top conditional sums (dnx, dnz, j1)
  double sa, sb;                // top-band signs
  int bot0, bot1, top0, top1;   // wall band plane ranges
  int jref_bot, jref_top;       // j1 = j - jref_bot, jref_top - j
};
void launch_spectra(Solver& s, const StatsSpecArgs& A);
void launch_omega_x(Solver& s, const StatsWxArgs& A, int nplanes);
// plane sums of chunk planes 0..nplanes-1 into profile entries j0.. ; tile
// sums of the chunk planes whose plane j (jfirst + t) lies in a wall band
void launch_stat_sums(Solver& s, const StatSumTask* tasks, int ntasks, int nplanes, int j0);
void launch_stat_cond(Solver& s, const StatCondTask* tasks, int ntasks, int nplanes, int dnx, int dnz, int jfirst);
// launchers (kernels_fused.hip)
void launch_ib_rhs(Solver& s, int kRK);
void launch_rhs_solve(Solver& s, int kRK);
void launch_pressure(Solver& s, int kRK);
void launch_meanflow_lds(Solver& s, int kRK);
void launch_mf_u11(Solver& s, int kRK, hipStream_t st);  // u11 and Q(u11) into d_u11
void launch_mf_apply(Solver& s);                       // g from Q(u), u += g u11
// multi-band resolution (kernels_band.hip)
void launch_band_gather(Solver& s);   // long columns of u1, u2, u3 -> the central solver's fields
// the same over planes [q0, q1) of every field (q1 at most the largest plane
// count; each field stops at its own), p too with with_p
void launch_band_gather(Solver& s, int q0, int q1, bool with_p);
void launch_band_scatter(Solver& s);  // central planes of the nonlinear scratch <- the central solver
void launch_cfl_bands(Solver& s);     // dt from the wall bands' and the central band's maxima
// Batched launches over the members of an ensemble: one launch of a kernel
// class covers the blocks of several solvers that use the same kernel
// variant.  Member i owns the linear blocks [off[i], off[i] + cnt[i]) (off
// a multiple of 8, so a member's block b runs on XCD b % 8 as in its own
// launch); its local block lb = b - off[i] is (lb % gx[i], lb / gx[i]).
// Blocks in the padding between members leave at once.
constexpr int kMaxBatch = 32;
struct BMap {
  int n;
  int off[kMaxBatch];
  int cnt[kMaxBatch];
  int gx[kMaxBatch];
};
__device__ __forceinline__ int bmap_member(const BMap& m, int b, int& bx, int& by) {
  int i = 0;
  while (i + 1 < m.n && b >= m.off[i + 1]) ++i;
  const int lb = b - m.off[i];
  if (lb >= m.cnt[i]) return -1;
  by = lb / m.gx[i];
  bx = lb - by * m.gx[i];
  return i;
}
// Kernel classes with batched launches.
enum BatchClass { BC_ZINV_NL = 0, BC_XPASS, BC_ZFWD, BC_RHS, BC_PRESS, BC_N };
// One member's part of a batched launch: `key` names the kernel variant,
// args the bytes of its argument struct (copied to the device before
// capture), grid its own block count (gx x gy), block its shape and
// lds_fixed + lds_thread x (threads per block) its dynamic LDS (a batch takes
// the largest block.y and lds_fixed of its parts; the kernels idle the
// thread rows a member does not use).
struct BatchPart {
  int key = -1;
  int gx = 0, gy = 0;
  dim3 block;
  size_t lds_fixed = 0, lds_thread = 0;
  std::vector<char> args;
  const void* tw = nullptr;  // twiddle table of the transform length
};
// A batch ready to launch: the parts' arguments at dargs on the device.
struct BatchLaunch {
  int cls = 0, key = -1;
  int lead = 0;             // index of the batch's first member (its stream carries the launch)
  std::vector<int> members;  // indices of the members in the batch
  dim3 grid, block;
  size_t lds = 0;
  BMap map;
  const void* dargs = nullptr;
  const void* tw = nullptr;
};
// class cls of solver s at substep kRK as a batch part; false: the solver
// has no batched form of the class (another engine or shape), message in err
bool batch_part_fft(Solver& s, int cls, int kRK, BatchPart& part, std::string& err);
bool batch_part_fused(Solver& s, int cls, int kRK, BatchPart& part, std::string& err);
void batch_launch_fft(hipStream_t st, const BatchLaunch& b);
void batch_launch_fused(hipStream_t st, const BatchLaunch& b);

// no-allocation physical planes (noalloc.hip): the file-scope scratch's
// device address is taken once per device (noalloc_init, device current);
// four_to_phys_noalloc queues the planes [plane0, plane0 + nplanes) of field
// `which` into the device buffer phys on the solver's stream
int noalloc_init(int dev, std::string& err);
int four_to_phys_noalloc(Solver& s, int which, int plane0, int nplanes, double* phys, std::string& err);

// device complex helpers
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscl(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// (0,k) * u
__host__ __device__ inline double2 imulk(double k, double2 u) { return make_double2(-k * u.y, k * u.x); }

}  // namespace cfd
