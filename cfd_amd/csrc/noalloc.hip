// noalloc.hip -- the physical planes of a field without any device allocation.
//
// cfd_four_to_phys allocates its half-transformed planes and job tables per
// call.  Callers that cannot allocate (code under HIP graph capture, tight
// serving loops) use cfd_four_to_phys_planes instead: the same z-inverse and
// paired x-inverse kernels, with the half-transformed planes and the job
// tables in one file-scope device array of kScratchBytes per GPU, and the
// physical planes written straight into the caller's device buffer.  Planes
// are paired for the x-inverse by their index in the field, as
// cfd_four_to_phys pairs them, so both give the same values bit for bit.
//
// The array's host-side name is not a device address: its address on each
// device comes from hipGetSymbolAddress, taken once per device with that
// device current (noalloc_init, called by cfd_create) and kept in
// g_scratch[device]; kernels and launches use only that pointer.  Every call
// on one GPU shares the array, so calls that overlap on one GPU (two handles
// on different streams, or one handle used from two threads) race.  Nothing in
// it needs to start at zero: the job tables are written by k_f2p_jobs, every
// element of a half-transformed row that the x-inverse reads is written by the
// z-inverse of the same call first, and the dump plane (the output of a pair
// partner outside the requested range) is written, never read.
#include <algorithm>
#include <mutex>

#include "common.hpp"

namespace cfd {

constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB per GPU
constexpr int kMaxDevices = 64;

namespace {

__device__ __align__(256) unsigned char scratch[kScratchBytes];

void* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

// job tables of one call.  The x-inverse pairs planes (2m, 2m+1) of the
// field, as cfd_four_to_phys pairs them, so that every plane gets the same
// rounding whatever range is asked for: the planes transformed are [e0, e0 +
// ne) with e0 = plane0 rounded down to even and the end rounded up to even
// (or the field's end); z-inverse job q takes plane e0 + q to half-transformed
// plane q; x job m pairs half-transformed planes 2m and 2m+1 (the field's
// last plane pairs with itself when the field has an odd count) and writes
// each plane of the requested range to phys, a partner outside it to `dump`
__global__ void k_f2p_jobs(ZInvJob* zj, XJob* xj, const double2* src, double2* z, double* phys, double* dump,
                           int plane0, int nplanes, int e0, int ne, size_t zpl, size_t ppl) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < ne) {
    ZInvJob j{};
    j.src = src;
    j.dst = z;
    j.pa = e0 + q;
    j.pb = 0;
    j.pc = 0;
    j.pd = q;
    j.mode = 0;
    j.flags = 1;  // leave out kz = -N2/2 as four_to_phys's resize does
    j.wa = j.wb = j.g = 0.0;
    zj[q] = j;
  }
  if (q < (ne + 1) / 2) {
    const int p = 2 * q;
    const bool pair = p + 1 < ne;
    auto out = [&](int plane) {
      const int o = e0 + plane - plane0;
      return (o >= 0 && o < nplanes) ? phys + (size_t)o * ppl : dump;
    };
    XJob j{};
    j.za = z + (size_t)p * zpl;
    j.zb = pair ? z + (size_t)(p + 1) * zpl : j.za;
    j.pa = out(p);
    j.pb = pair ? out(p + 1) : j.pa;
    j.zo = nullptr;
    j.plane = 0;
    j.pad = 0;
    xj[q] = j;
  }
}

}  // namespace

int noalloc_init(int dev, std::string& err) {
  if (dev < 0 || dev >= kMaxDevices) {
    err = "device ordinal out of range for the no-allocation scratch";
    return CFD_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[dev]) return CFD_OK;
  // the caller has made dev current
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return CFD_EHIP;
  }
  g_scratch[dev] = ptr;
  return CFD_OK;
}

int four_to_phys_noalloc(Solver& s, int which, int plane0, int nplanes, double* phys, std::string& err) {
  if (s.dev < 0 || s.dev >= kMaxDevices || !g_scratch[s.dev]) {
    err = "no-allocation scratch not set up on this device";
    return CFD_EINVAL;
  }
  const size_t zpl = (size_t)s.ngz * s.Mx, ppl = (size_t)s.ngz * s.ngx;
  const int ny = which == CFD_U2 || which == CFD_NU2 || which == CFD_DU2
                     ? s.nyv
                     : (which == CFD_P || which == CFD_PSI || which == CFD_DIV ? s.nyp : s.nyu);
  const int e0 = plane0 & ~1, e1 = std::min(ny, (plane0 + nplanes + 1) & ~1), ne = e1 - e0;
  const int nx = (ne + 1) / 2;
  const size_t zj_bytes = up256(sizeof(ZInvJob) * (size_t)ne);
  const size_t xj_bytes = up256(sizeof(XJob) * (size_t)nx);
  const size_t z_bytes = up256(sizeof(double2) * zpl * (size_t)ne);
  const size_t need = zj_bytes + xj_bytes + z_bytes + sizeof(double) * ppl;
  if (need > kScratchBytes) {
    const size_t per = sizeof(double2) * zpl + sizeof(ZInvJob) + sizeof(XJob);
    const long long room = ((long long)kScratchBytes - (long long)(sizeof(double) * ppl + 512)) / (long long)per - 2;
    err = "four_to_phys_planes: " + std::to_string(nplanes) + " planes need " + std::to_string(need) +
          " bytes of scratch, more than the " + std::to_string(kScratchBytes) + " bytes per GPU (at most " +
          std::to_string(std::max(room, 0LL)) + " planes of this grid per call)";
    return CFD_EINVAL;
  }
  unsigned char* base = static_cast<unsigned char*>(g_scratch[s.dev]);
  ZInvJob* zj = reinterpret_cast<ZInvJob*>(base);
  XJob* xj = reinterpret_cast<XJob*>(base + zj_bytes);
  double2* z = reinterpret_cast<double2*>(base + zj_bytes + xj_bytes);
  double* dump = reinterpret_cast<double*>(base + zj_bytes + xj_bytes + z_bytes);
  const int nb = (ne + 255) / 256;
  hipLaunchKernelGGL(k_f2p_jobs, dim3(nb), dim3(256), 0, s.stream, zj, xj, (const double2*)s.F[which], z, phys, dump,
                     plane0, nplanes, e0, ne, zpl, ppl);
  launch_zinv(s, zj, ne, false);
  launch_xpass_inv2(s, xj, nx);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    err = std::string("four_to_phys_planes: ") + hipGetErrorString(e);
    return CFD_EHIP;
  }
  return CFD_OK;
}

}  // namespace cfd
