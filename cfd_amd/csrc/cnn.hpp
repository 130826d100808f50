// cnn.hpp -- the closure network object and its launchers (cnn.cpp,
// kernels_cnn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <string>
#include <vector>

#include "cfd_amd.h"

namespace cfd {

constexpr int kLayers = 4;
constexpr int kLayerSize = 27 * 32 * 32 + 32;  // padded W[t][ci][co] + b[co] of one layer
constexpr int kLossBlocks = 256;
extern const int kCinReal[kLayers], kCoutReal[kLayers];

}  // namespace cfd

struct cfd_cnn_s {
  int dev = 0;
  hipStream_t stream = nullptr;
  int NY = 0, NZ = 0, NX = 0;
  size_t npts = 0;
  int nblk = 0;  // blocks of the weight-gradient kernel
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;  // FP32 master, gradient, Adam moments
  float* partial = nullptr;                                        // weight-gradient partials (nblk layers)
  float* lsum = nullptr;
  __bf16 *wf[cfd::kLayers] = {}, *wb[cfd::kLayers] = {};          // forward / backward-data operands
  __bf16 *x = nullptr, *t = nullptr;                               // input [p][4], target [p][4]
  __bf16 *z[3] = {}, *a[3] = {}, *d[2] = {};                       // pre-activations, activations, gradients
  float* y = nullptr;                                              // output [p][3]
  __bf16 *dx = nullptr, *dt = nullptr;                             // dataset slots
  int nslots = 0, adam_step = 0;
  // normalisation of the closure data: inputs u1, u2, u3 divided by
  // scale[0..2], targets (corrections) by scale[3..5]; 0 = not set yet
  float scale[6] = {0, 0, 0, 0, 0, 0};
  double bytes = 0.0;
  std::vector<void*> allocs;
  std::string err;
};

namespace cfd {

int cnn_nparams();
int cnn_relayout(cfd_cnn_s* c);
int cnn_forward_dev(cfd_cnn_s* c, const __bf16* x);
int cnn_backward_dev(cfd_cnn_s* c, const __bf16* x, const __bf16* tgt, int tc);
int cnn_adam_dev(cfd_cnn_s* c, float lr, float b1, float b2, float eps);

// kernels_cnn.hip
bool cnn_dims_ok(int NY, int NZ, int NX);
void launch_conv_fwd(hipStream_t st, int layer, const __bf16* in, const __bf16* wf, const float* bias, __bf16* zout,
                     __bf16* aout, float* yout, int nout, int NY, int NZ, int NX);
void launch_conv_bwd_data(hipStream_t st, const __bf16* dz, const __bf16* wb, const __bf16* zprev, __bf16* dprev,
                          int NY, int NZ, int NX);
void launch_conv_dw(hipStream_t st, int layer, const __bf16* in, const __bf16* dz, float* partial, int nblk,
                    float* grad, int NY, int NZ, int NX);
void launch_loss(hipStream_t st, const float* y, const __bf16* tgt, int tc, int nout, size_t npts, __bf16* dy,
                 float* lsum, int nblk);
void launch_adam(hipStream_t st, float* w, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2,
                 float eps, int step);
void launch_wlayouts(hipStream_t st, const float* w, int cin, __bf16* wf, __bf16* wb);

}  // namespace cfd

namespace cfd {
// closure data movement (kernels_cnn.hip)
void launch_filter(hipStream_t st, const double2* fine, const double2* other, double2* out, int Mxf, int N2f, int Mxc,
                   int N2c, int ngxc, int ngzc, double a, double b, int nplanes);
void launch_pack(hipStream_t st, const double* p1, const double* p2, const double* p3, const unsigned char* mask,
                 int nyu, int nyv, int ngz, int ngx, const float* s, int withmask, __bf16* out);
// u1..u3 null: the corrections as predicted; else backscatter-clipped
void launch_unpack(hipStream_t st, const float* y, int nyu, int nyv, int ngz, int ngx, const float* s, double* p1,
                   double* p2, double* p3, const double* u1, const double* u2, const double* u3);
void launch_add_planes(hipStream_t st, double2* a, const double2* c, int ncol, int q0, int q1);
void launch_add_planes(hipStream_t st, float2* a, const double2* c, int ncol, int q0, int q1);
void launch_zero_net_force(hipStream_t st, double2* c, int ncol, int q0, int q1, const double* yv);
void launch_sumsq(hipStream_t st, const double* v, size_t n, double* acc);
}  // namespace cfd
