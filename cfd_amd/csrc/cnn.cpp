// cnn.cpp -- the closure network object (include/cfd_amd.h, cfd_cnn_*):
// parameters, activations, training steps, and a device dataset of
// (input, target) volumes.  Kernels: kernels_cnn.hip.
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cnn.hpp"

namespace cfd {

namespace {
thread_local std::string g_cnn_err;

int cfail(cfd_cnn_s* c, int rc, const std::string& msg) {
  g_cnn_err = msg;
  if (c) c->err = msg;
  return rc;
}
#define CHK(c, x)                                                                                    \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return cfail((c), CFD_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int alloc(cfd_cnn_s* c, T** p, size_t n) {
  void* q = nullptr;
  if (hipMalloc(&q, (n ? n : 1) * sizeof(T)) != hipSuccess) return cfail(c, CFD_ENOMEM, "hipMalloc failed");
  (void)hipMemset(q, 0, (n ? n : 1) * sizeof(T));
  c->allocs.push_back(q);
  c->bytes += (double)(n ? n : 1) * sizeof(T);
  *p = static_cast<T*>(q);
  return 0;
}

}  // namespace

const int kCinReal[kLayers] = {4, 32, 32, 32};
const int kCoutReal[kLayers] = {32, 32, 32, 3};

int cnn_nparams() {
  int n = 0;
  for (int l = 0; l < kLayers; ++l) n += 27 * kCinReal[l] * kCoutReal[l] + kCoutReal[l];
  return n;
}

int cnn_relayout(cfd_cnn_s* c) {
  for (int l = 0; l < kLayers; ++l)
    launch_wlayouts(c->stream, c->P + (size_t)l * kLayerSize, l == 0 ? 16 : 32, c->wf[l], c->wb[l]);
  CHK(c, hipGetLastError());
  return 0;
}

int cnn_forward_dev(cfd_cnn_s* c, const __bf16* x) {
  const int NY = c->NY, NZ = c->NZ, NX = c->NX;
  auto bias = [&](int l) { return c->P + (size_t)l * kLayerSize + 27 * 32 * 32; };
  launch_conv_fwd(c->stream, 0, x, c->wf[0], bias(0), c->z[0], c->a[0], nullptr, 0, NY, NZ, NX);
  launch_conv_fwd(c->stream, 1, c->a[0], c->wf[1], bias(1), c->z[1], c->a[1], nullptr, 0, NY, NZ, NX);
  launch_conv_fwd(c->stream, 2, c->a[1], c->wf[2], bias(2), c->z[2], c->a[2], nullptr, 0, NY, NZ, NX);
  launch_conv_fwd(c->stream, 3, c->a[2], c->wf[3], bias(3), nullptr, nullptr, c->y, 3, NY, NZ, NX);
  CHK(c, hipGetLastError());
  return 0;
}

// gradients of the mean squared error against tgt (tc channels per point, the
// first 3 used) into c->G; the forward pass of x must be the last one run
int cnn_backward_dev(cfd_cnn_s* c, const __bf16* x, const __bf16* tgt, int tc) {
  const int NY = c->NY, NZ = c->NZ, NX = c->NX;
  launch_loss(c->stream, c->y, tgt, tc, 3, c->npts, c->d[0], c->lsum, kLossBlocks);
  auto g = [&](int l) { return c->G + (size_t)l * kLayerSize; };
  // layer 3 (linear): dW from a[2]; gradient into the output of layer 2
  launch_conv_dw(c->stream, 3, c->a[2], c->d[0], c->partial, c->nblk, g(3), NY, NZ, NX);
  launch_conv_bwd_data(c->stream, c->d[0], c->wb[3], c->z[2], c->d[1], NY, NZ, NX);
  launch_conv_dw(c->stream, 2, c->a[1], c->d[1], c->partial, c->nblk, g(2), NY, NZ, NX);
  launch_conv_bwd_data(c->stream, c->d[1], c->wb[2], c->z[1], c->d[0], NY, NZ, NX);
  launch_conv_dw(c->stream, 1, c->a[0], c->d[0], c->partial, c->nblk, g(1), NY, NZ, NX);
  launch_conv_bwd_data(c->stream, c->d[0], c->wb[1], c->z[0], c->d[1], NY, NZ, NX);
  launch_conv_dw(c->stream, 0, x, c->d[1], c->partial, c->nblk, g(0), NY, NZ, NX);
  CHK(c, hipGetLastError());
  return 0;
}

int cnn_adam_dev(cfd_cnn_s* c, float lr, float b1, float b2, float eps) {
  c->adam_step += 1;
  launch_adam(c->stream, c->P, c->G, c->M1, c->M2, kLayers * kLayerSize, lr, b1, b2, eps, c->adam_step);
  CHK(c, hipGetLastError());
  return cnn_relayout(c);
}

int cnn_loss_read(cfd_cnn_s* c, double* loss) {
  std::vector<float> ls(kLossBlocks);
  CHK(c, hipStreamSynchronize(c->stream));
  CHK(c, hipMemcpy(ls.data(), c->lsum, ls.size() * sizeof(float), hipMemcpyDeviceToHost));
  double s = 0.0;
  for (float v : ls) s += v;
  *loss = s;
  return 0;
}

}  // namespace cfd

using namespace cfd;

extern "C" int cfd_cnn_nparams(void) { return cnn_nparams(); }

extern "C" int cfd_cnn_create(int device, int ny, int nz, int nx, cfd_cnn* out) {
  if (!out) return cfail(nullptr, CFD_EINVAL, "null argument");
  *out = nullptr;
  if (!cnn_dims_ok(ny, nz, nx))
    return cfail(nullptr, CFD_EINVAL, "volume: nx must be a multiple of 32 and nz of 8");
  cfd_cnn_s* c = new cfd_cnn_s();
  auto bail = [&](int rc) {
    g_cnn_err = c->err.empty() ? g_cnn_err : c->err;
    cfd_cnn_destroy(c);
    return rc;
  };
  c->dev = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    c->err = "device or stream";
    return bail(CFD_EHIP);
  }
  c->NY = ny;
  c->NZ = nz;
  c->NX = nx;
  c->npts = (size_t)ny * nz * nx;
  const int tiles = (nx / 32) * (nz / 4) * ny;
  c->nblk = tiles < 1024 ? tiles : 1024;
  int rc;
  const size_t np = (size_t)kLayers * kLayerSize;
  if ((rc = alloc(c, &c->P, np)) || (rc = alloc(c, &c->G, np)) || (rc = alloc(c, &c->M1, np)) ||
      (rc = alloc(c, &c->M2, np)) || (rc = alloc(c, &c->partial, (size_t)c->nblk * kLayerSize)) ||
      (rc = alloc(c, &c->lsum, kLossBlocks)) || (rc = alloc(c, &c->x, c->npts * 4)) ||
      (rc = alloc(c, &c->t, c->npts * 4)) || (rc = alloc(c, &c->y, c->npts * 3)))
    return bail(rc);
  for (int l = 0; l < kLayers; ++l)
    if ((rc = alloc(c, &c->wf[l], (size_t)27 * 32 * (l == 0 ? 16 : 32))) || (rc = alloc(c, &c->wb[l], 27 * 32 * 32)))
      return bail(rc);
  for (int l = 0; l < 3; ++l)
    if ((rc = alloc(c, &c->z[l], c->npts * 32)) || (rc = alloc(c, &c->a[l], c->npts * 32))) return bail(rc);
  for (int l = 0; l < 2; ++l)
    if ((rc = alloc(c, &c->d[l], c->npts * 32))) return bail(rc);
  *out = c;
  return CFD_OK;
}

extern "C" int cfd_cnn_destroy(cfd_cnn c) {
  if (!c) return CFD_OK;
  hipSetDevice(c->dev);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (void* q : c->allocs) hipFree(q);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return CFD_OK;
}

extern "C" const char* cfd_cnn_last_error(cfd_cnn c) { return c ? c->err.c_str() : g_cnn_err.c_str(); }

// real parameters <-> padded device layout: per layer W[t][ci][co] (27 x cin x
// cout of the real sizes) then b[co]
extern "C" int cfd_cnn_set_params(cfd_cnn c, const float* w) {
  if (!c || !w) return cfail(c, CFD_EINVAL, "null argument");
  std::vector<float> P((size_t)kLayers * kLayerSize, 0.0f);
  size_t o = 0;
  for (int l = 0; l < kLayers; ++l) {
    float* L = P.data() + (size_t)l * kLayerSize;
    for (int t = 0; t < 27; ++t)
      for (int ci = 0; ci < kCinReal[l]; ++ci)
        for (int co = 0; co < kCoutReal[l]; ++co) L[((size_t)t * 32 + ci) * 32 + co] = w[o++];
    for (int co = 0; co < kCoutReal[l]; ++co) L[27 * 32 * 32 + co] = w[o++];
  }
  hipSetDevice(c->dev);
  CHK(c, hipStreamSynchronize(c->stream));
  CHK(c, hipMemcpy(c->P, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice));
  CHK(c, hipMemset(c->M1, 0, P.size() * sizeof(float)));
  CHK(c, hipMemset(c->M2, 0, P.size() * sizeof(float)));
  c->adam_step = 0;
  return cnn_relayout(c);
}

static int gather(cfd_cnn c, const float* dev, float* w) {
  std::vector<float> P((size_t)kLayers * kLayerSize);
  hipSetDevice(c->dev);
  CHK(c, hipStreamSynchronize(c->stream));
  CHK(c, hipMemcpy(P.data(), dev, P.size() * sizeof(float), hipMemcpyDeviceToHost));
  size_t o = 0;
  for (int l = 0; l < kLayers; ++l) {
    const float* L = P.data() + (size_t)l * kLayerSize;
    for (int t = 0; t < 27; ++t)
      for (int ci = 0; ci < kCinReal[l]; ++ci)
        for (int co = 0; co < kCoutReal[l]; ++co) w[o++] = L[((size_t)t * 32 + ci) * 32 + co];
    for (int co = 0; co < kCoutReal[l]; ++co) w[o++] = L[27 * 32 * 32 + co];
  }
  return CFD_OK;
}

extern "C" int cfd_cnn_get_params(cfd_cnn c, float* w) {
  if (!c || !w) return cfail(c, CFD_EINVAL, "null argument");
  return gather(c, c->P, w);
}
extern "C" int cfd_cnn_get_grads(cfd_cnn c, float* g) {
  if (!c || !g) return cfail(c, CFD_EINVAL, "null argument");
  return gather(c, c->G, g);
}

// host volumes: x[p][4] (u, v, w, mask), target[p][3], output y[p][3]
static int upload_bf16(cfd_cnn c, __bf16* dev, const float* h, int hc, int dc) {
  std::vector<__bf16> b(c->npts * dc);
  for (size_t p = 0; p < c->npts; ++p)
    for (int k = 0; k < dc; ++k) b[p * dc + k] = (__bf16)(k < hc ? h[p * hc + k] : 0.0f);
  hipSetDevice(c->dev);
  CHK(c, hipStreamSynchronize(c->stream));
  CHK(c, hipMemcpy(dev, b.data(), b.size() * sizeof(__bf16), hipMemcpyHostToDevice));
  return CFD_OK;
}
extern "C" int cfd_cnn_set_input(cfd_cnn c, const float* x) {
  if (!c || !x) return cfail(c, CFD_EINVAL, "null argument");
  return upload_bf16(c, c->x, x, 4, 4);
}
extern "C" int cfd_cnn_set_target(cfd_cnn c, const float* t) {
  if (!c || !t) return cfail(c, CFD_EINVAL, "null argument");
  return upload_bf16(c, c->t, t, 3, 4);
}
extern "C" int cfd_cnn_forward(cfd_cnn c) {
  if (!c) return cfail(c, CFD_EINVAL, "null handle");
  hipSetDevice(c->dev);
  return cnn_forward_dev(c, c->x);
}
extern "C" int cfd_cnn_get_output(cfd_cnn c, float* y) {
  if (!c || !y) return cfail(c, CFD_EINVAL, "null argument");
  hipSetDevice(c->dev);
  CHK(c, hipStreamSynchronize(c->stream));
  CHK(c, hipMemcpy(y, c->y, c->npts * 3 * sizeof(float), hipMemcpyDeviceToHost));
  return CFD_OK;
}
extern "C" int cfd_cnn_backward(cfd_cnn c, double* loss) {
  if (!c) return cfail(c, CFD_EINVAL, "null handle");
  hipSetDevice(c->dev);
  if (int rc = cnn_backward_dev(c, c->x, c->t, 4)) return rc;
  return loss ? cnn_loss_read(c, loss) : CFD_OK;
}
extern "C" int cfd_cnn_adam(cfd_cnn c, float lr, float b1, float b2, float eps) {
  if (!c) return cfail(c, CFD_EINVAL, "null handle");
  hipSetDevice(c->dev);
  return cnn_adam_dev(c, lr, b1, b2, eps);
}

// the device dataset: nslots (input, target) volume pairs
extern "C" int cfd_cnn_dataset(cfd_cnn c, int nslots) {
  if (!c || nslots < 1) return cfail(c, CFD_EINVAL, "nslots must be >= 1");
  if (c->nslots) return cfail(c, CFD_EINVAL, "dataset already allocated");
  hipSetDevice(c->dev);
  if (int rc = alloc(c, &c->dx, c->npts * 4 * (size_t)nslots)) return rc;
  if (int rc = alloc(c, &c->dt, c->npts * 4 * (size_t)nslots)) return rc;
  c->nslots = nslots;
  return CFD_OK;
}

// one training step on dataset slot s: forward, loss, gradients, Adam
extern "C" int cfd_cnn_train_slot(cfd_cnn c, int slot, float lr, float b1, float b2, float eps, double* loss) {
  if (!c || slot < 0 || slot >= c->nslots) return cfail(c, CFD_EINVAL, "bad slot");
  hipSetDevice(c->dev);
  const __bf16* x = c->dx + c->npts * 4 * (size_t)slot;
  const __bf16* t = c->dt + c->npts * 4 * (size_t)slot;
  if (int rc = cnn_forward_dev(c, x)) return rc;
  if (int rc = cnn_backward_dev(c, x, t, 4)) return rc;
  if (int rc = cnn_adam_dev(c, lr, b1, b2, eps)) return rc;
  return loss ? cnn_loss_read(c, loss) : CFD_OK;
}

// loss of the current parameters on slot s (no update)
extern "C" int cfd_cnn_eval_slot(cfd_cnn c, int slot, double* loss) {
  if (!c || !loss || slot < 0 || slot >= c->nslots) return cfail(c, CFD_EINVAL, "bad slot");
  hipSetDevice(c->dev);
  const __bf16* x = c->dx + c->npts * 4 * (size_t)slot;
  const __bf16* t = c->dt + c->npts * 4 * (size_t)slot;
  if (int rc = cnn_forward_dev(c, x)) return rc;
  launch_loss(c->stream, c->y, t, 4, 3, c->npts, c->d[0], c->lsum, kLossBlocks);
  return cnn_loss_read(c, loss);
}

// copy a dataset slot to the network input/target (host-visible tests)
extern "C" int cfd_cnn_load_slot(cfd_cnn c, int slot) {
  if (!c || slot < 0 || slot >= c->nslots) return cfail(c, CFD_EINVAL, "bad slot");
  hipSetDevice(c->dev);
  CHK(c, hipMemcpyAsync(c->x, c->dx + c->npts * 4 * (size_t)slot, c->npts * 4 * sizeof(__bf16),
                        hipMemcpyDeviceToDevice, c->stream));
  CHK(c, hipMemcpyAsync(c->t, c->dt + c->npts * 4 * (size_t)slot, c->npts * 4 * sizeof(__bf16),
                        hipMemcpyDeviceToDevice, c->stream));
  CHK(c, hipStreamSynchronize(c->stream));
  return CFD_OK;
}

extern "C" int cfd_cnn_device_bytes(cfd_cnn c, double* bytes) {
  if (!c || !bytes) return cfail(c, CFD_EINVAL, "null argument");
  *bytes = c->bytes;
  return CFD_OK;
}

extern "C" int cfd_cnn_set_scales(cfd_cnn c, const float* s6) {
  if (!c || !s6) return cfail(c, CFD_EINVAL, "null argument");
  for (int k = 0; k < 6; ++k)
    if (!(s6[k] > 0.0f)) return cfail(c, CFD_EINVAL, "scales must be positive");
  for (int k = 0; k < 6; ++k) c->scale[k] = s6[k];
  return CFD_OK;
}
extern "C" int cfd_cnn_get_scales(cfd_cnn c, float* s6) {
  if (!c || !s6) return cfail(c, CFD_EINVAL, "null argument");
  for (int k = 0; k < 6; ++k) s6[k] = c->scale[k];
  return CFD_OK;
}
