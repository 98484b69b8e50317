// sx_learn.cpp -- the U-Net of the learned coarse-grid preconditioner:
// parameters, activations, forward and backward passes and Adam, over the
// kernels of sx_unet.hip, plus the C-ABI entry points.
//
// Architecture (channels-last tensors on vertex-centred grids L0..L3):
//   L0: a0 = relu(conv(x, 3 -> 16)),  b0 = relu(conv(a0, 16 -> 16))
//   L1: a1 = relu(conv_s2(b0, 16 -> 32)), b1 = relu(conv(a1, 32 -> 32))
//   L2: a2 = relu(conv_s2(b1, 32 -> 64)), b2 = relu(conv(a2, 64 -> 64))
//   L3: a3 = relu(conv_s2(b2, 64 -> 64)), b3 = relu(conv(a3, 64 -> 64))
//   L2: c2 = relu(conv([P b3, b2], 128 -> 64))
//   L1: c1 = relu(conv([P c2, b1], 96 -> 32))
//   L0: c0 = relu(conv([P c1, b0], 48 -> 16))
//   y  = conv(c0, 16 -> 3)          (FP32 output)
// P is the vertex-centred trilinear prolongation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/serac_amd.h"
#include "sx_unet.hpp"

namespace sxh {
void set_last_error(const std::string& m);  // sx_capi.cpp
}

namespace {

struct unet_failure : std::runtime_error {
  int code;
  unet_failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define UN_HIP(expr)                                                                                     \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess) throw unet_failure(SX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Buf {
  void*  p     = nullptr;
  size_t bytes = 0;
  Buf()        = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf()
  {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t n)
  {
    if (p) (void)hipFree(p);
    p     = nullptr;
    bytes = n;
    if (n) UN_HIP(hipMalloc(&p, n));
  }
  template <typename T>
  T* as() const
  {
    return static_cast<T*>(p);
  }
};

struct LayerDef {
  int cin_a, cin_b, cout, mode, relu, lin, lout;  // lin / lout: grid level of input / output
};

// a0 b0 a1 b1 a2 b2 a3 b3 c2 c1 c0 y
const LayerDef kLayers[12] = {
    {3, 0, 16, 0, 1, 0, 0},   {16, 0, 16, 0, 1, 0, 0},  {16, 0, 32, 1, 1, 0, 1},  {32, 0, 32, 0, 1, 1, 1},
    {32, 0, 64, 1, 1, 1, 2},  {64, 0, 64, 0, 1, 2, 2},  {64, 0, 64, 1, 1, 2, 3},  {64, 0, 64, 0, 1, 3, 3},
    {64, 64, 64, 0, 1, 2, 2}, {64, 32, 32, 0, 1, 1, 1}, {32, 16, 16, 0, 1, 0, 0}, {16, 0, 3, 0, 0, 0, 0},
};

}  // namespace

struct sx_unet {
  sxu::Dims3  dims[4];
  int64_t     nvox[4];
  hipStream_t stream = nullptr;
  // parameters: per layer weights [27][cin][cout] then bias [cout], FP32 master
  int64_t woff[12], boff[12], nparam = 0;
  Buf     P, G, M, V, Wbf;  // FP32 master, gradient, Adam moments; BF16 weights (same offsets)
  int     step     = 0;
  int     identity = 0;  // 1: no ReLU (a linear network)
  // BF16 weights k-contiguous per output channel for the MFMA convolutions,
  // rebuilt from Wbf on the first forward pass after the weights change
  Buf     Wt;
  int64_t wtoff[12] = {};
  bool    wt_dirty  = true;
  int     use_mfma  = 1;
  // activations (BF16) and their gradients (FP32)
  Buf act[12];   // output of layer k
  Buf up[3];     // prolonged inputs of layers 8, 9, 10
  Buf gact[12];  // dL/d(output of layer k)
  Buf gup[3];
  Buf xin;       // BF16 input [n0][3]
  Buf yout;      // FP32 output [n0][3]
  Buf wflip, tmp, sums;

  int64_t cin(int k) const { return kLayers[k].cin_a + kLayers[k].cin_b; }
};

namespace {

template <typename F>
int unet_api(F&& f)
{
  try {
    f();
    return SX_OK;
  } catch (const unet_failure& e) {
    sxh::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    sxh::set_last_error(e.what());
    return SX_ERR_ARG;
  }
}

sxu::ConvArgs conv_args(const sx_unet* N, int k, const void* xa, const void* xb)
{
  const LayerDef& L = kLayers[k];
  sxu::ConvArgs   A{};
  A.xa = xa;
  A.ca = L.cin_a;
  A.xb = xb;
  A.cb = L.cin_b;
  for (int d = 0; d < 3; d++) {
    A.din[d]  = N->dims[L.lin].n[d];
    A.dout[d] = N->dims[L.lout].n[d];
  }
  A.mode = L.mode;
  A.w    = N->Wbf.as<uint16_t>() + N->woff[k];
  A.bias = N->P.as<float>() + N->boff[k];
  A.relu = L.relu && !N->identity;
  return A;
}

void forward(sx_unet* N)
{
  hipStream_t s = N->stream;
  if (N->use_mfma && N->wt_dirty) {
    for (int k = 0; k < 12; k++)
      if (N->wtoff[k] >= 0)
        UN_HIP(sxu::weights_kmajor(N->Wbf.as<uint16_t>() + N->woff[k], int(N->cin(k)), kLayers[k].cout,
                                   N->Wt.as<uint16_t>() + N->wtoff[k], s));
    N->wt_dirty = false;
  }
  auto run = [&](int k, const void* xa, const void* xb, void* y, bool out_f32) {
    sxu::ConvArgs A = conv_args(N, k, xa, xb);
    A.y             = y;
    if (N->use_mfma && N->wtoff[k] >= 0 && sxu::conv_mfma_supported(A, kLayers[k].cout, false, out_f32))
      UN_HIP(sxu::conv_mfma(A, kLayers[k].cout, N->Wt.as<uint16_t>() + N->wtoff[k], s));
    else
      UN_HIP(sxu::conv(A, kLayers[k].cout, false, out_f32, s));
  };
  run(0, N->xin.p, nullptr, N->act[0].p, false);
  for (int k = 1; k < 8; k++) run(k, N->act[k - 1].p, nullptr, N->act[k].p, false);
  // decoder: layer 8 takes P b3 and b2, 9 takes P c2 and b1, 10 takes P c1 and b0
  const int src[3] = {7, 8, 9}, skip[3] = {5, 3, 1};
  for (int j = 0; j < 3; j++) {
    const int k = 8 + j, lc = kLayers[src[j]].lout, lf = kLayers[k].lout;
    UN_HIP(sxu::prolong_bf16(N->act[src[j]].as<uint16_t>(), N->dims[lc], N->up[j].as<uint16_t>(), N->dims[lf],
                             kLayers[k].cin_a, s));
    run(k, N->up[j].p, N->act[skip[j]].p, N->act[k].p, false);
  }
  run(11, N->act[10].p, nullptr, N->yout.p, true);
}

// data gradient of layer k's input part (c0, cpart) from g (FP32, dL/d pre-activation)
void data_grad(sx_unet* N, int k, const float* g, int c0, int cpart, float* out)
{
  const LayerDef& L = kLayers[k];
  hipStream_t     s = N->stream;
  UN_HIP(sxu::flip_weights(N->Wbf.as<uint16_t>() + N->woff[k], int(N->cin(k)), L.cout, c0, cpart,
                           N->wflip.as<uint16_t>(), s));
  sxu::ConvArgs A{};
  A.xa = g;
  A.ca = L.cout;
  A.xb = nullptr;
  A.cb = 0;
  for (int d = 0; d < 3; d++) {
    A.din[d]  = N->dims[L.lout].n[d];
    A.dout[d] = N->dims[L.lin].n[d];
  }
  A.mode = L.mode == 1 ? 2 : 0;
  A.w    = N->wflip.as<uint16_t>();
  A.bias = nullptr;
  A.relu = 0;
  A.y    = out;
  UN_HIP(sxu::conv(A, cpart, true, true, s));
}

void weight_grad(sx_unet* N, int k, const void* xa, const void* xb, const float* g)
{
  const LayerDef& L = kLayers[k];
  sxu::WGradArgs  A{};
  A.xa = xa;
  A.ca = L.cin_a;
  A.xb = xb;
  A.cb = L.cin_b;
  for (int d = 0; d < 3; d++) {
    A.din[d]  = N->dims[L.lin].n[d];
    A.dout[d] = N->dims[L.lout].n[d];
  }
  A.mode = L.mode;
  A.g    = g;
  A.dw   = N->G.as<float>() + N->woff[k];
  A.db   = N->G.as<float>() + N->boff[k];
  UN_HIP(sxu::conv_wgrad(A, L.cout, false, N->stream));
}

// backward from dL/dy (in gact[11]); gradients accumulate into G
void backward(sx_unet* N)
{
  hipStream_t s   = N->stream;
  auto        nel = [&](int k) { return N->nvox[kLayers[k].lout] * kLayers[k].cout; };
  float*      g;
  // head (no activation)
  g = N->gact[11].as<float>();
  weight_grad(N, 11, N->act[10].p, nullptr, g);
  data_grad(N, 11, g, 0, 16, N->gact[10].as<float>());
  // decoder 10, 9, 8: input parts [P c, skip]
  const int src[3] = {7, 8, 9}, skip[3] = {5, 3, 1};
  for (int j = 2; j >= 0; j--) {
    const int k = 8 + j;
    g           = N->gact[k].as<float>();
    if (!N->identity) UN_HIP(sxu::relu_bwd(g, N->act[k].as<uint16_t>(), nel(k), s));
    weight_grad(N, k, N->up[j].p, N->act[skip[j]].p, g);
    data_grad(N, k, g, 0, kLayers[k].cin_a, N->gup[j].as<float>());
    // the skip's gradient from the decoder starts the skip's accumulation
    data_grad(N, k, g, kLayers[k].cin_a, kLayers[k].cin_b, N->gact[skip[j]].as<float>());
    // the prolonged source's gradient is the restriction of gup
    const int lc = kLayers[src[j]].lout, lf = kLayers[k].lout;
    UN_HIP(sxu::restrict_f32(N->gup[j].as<float>(), N->dims[lf], N->gact[src[j]].as<float>(), N->dims[lc],
                             kLayers[k].cin_a, s));
  }
  // encoder 7 .. 0
  for (int k = 7; k >= 0; k--) {
    g = N->gact[k].as<float>();
    if (!N->identity) UN_HIP(sxu::relu_bwd(g, N->act[k].as<uint16_t>(), nel(k), s));
    const void* xin = k == 0 ? N->xin.p : N->act[k - 1].p;
    weight_grad(N, k, xin, nullptr, g);
    if (k == 0) break;
    // layer k's input is act[k-1]; odd encoder layers (b) feed the decoder
    // too, so their gradient already holds the skip term and gets this added
    float*        gin = N->gact[k - 1].as<float>();
    const int64_t nin = N->nvox[kLayers[k].lin] * kLayers[k].cin_a;
    const bool    has_skip = (k - 1 == 1 || k - 1 == 3 || k - 1 == 5);
    if (has_skip) {
      data_grad(N, k, g, 0, kLayers[k].cin_a, N->tmp.as<float>());
      UN_HIP(sxu::add_f32(gin, N->tmp.as<float>(), nin, s));
    } else {
      data_grad(N, k, g, 0, kLayers[k].cin_a, gin);
    }
  }
}

// Normal weights with variance gain / fan_in (fan_in = 27 cin), zero biases,
// cleared Adam state.  ReLU: gain 2 (He).  Identity: gain 1, and the output
// layer scaled by 1e-2 so that a new network starts near N = 0.
void init_params(sx_unet* N, uint64_t seed)
{
  const int64_t      off = N->nparam;
  std::vector<float> h(static_cast<size_t>(off), 0.0f);
  std::mt19937_64    rng(seed);
  for (int k = 0; k < 12; k++) {
    const float gain = N->identity ? 1.0f : 2.0f;
    const float sd   = std::sqrt(gain / float(27 * N->cin(k))) * (N->identity && k == 11 ? 0.01f : 1.0f);
    std::normal_distribution<float> nd(0.0f, sd);
    for (int64_t i = 0; i < 27 * N->cin(k) * kLayers[k].cout; i++) h[size_t(N->woff[k] + i)] = nd(rng);
  }
  UN_HIP(hipMemcpy(N->P.p, h.data(), size_t(off) * 4, hipMemcpyHostToDevice));
  UN_HIP(hipMemset(N->G.p, 0, size_t(off) * 4));
  UN_HIP(hipMemset(N->M.p, 0, size_t(off) * 4));
  UN_HIP(hipMemset(N->V.p, 0, size_t(off) * 4));
  UN_HIP(sxu::to_bf16(N->P.as<float>(), N->Wbf.as<uint16_t>(), off, nullptr));
  UN_HIP(hipDeviceSynchronize());
  N->step     = 0;
  N->wt_dirty = true;
}

void init(sx_unet* N, int nx, int ny, int nz, uint64_t seed)
{
  N->dims[0] = sxu::Dims3{{nx, ny, nz}};
  for (int l = 1; l < 4; l++)
    for (int d = 0; d < 3; d++) {
      const int n = N->dims[l - 1].n[d];
      if (n < 3 || (n - 1) % 2) throw unet_failure(SX_ERR_ARG, "vertex grid sizes must be 2^3 m + 1 (m >= 1)");
      N->dims[l].n[d] = (n - 1) / 2 + 1;
    }
  for (int l = 0; l < 4; l++) N->nvox[l] = int64_t(N->dims[l].n[0]) * N->dims[l].n[1] * N->dims[l].n[2];
  int64_t off = 0;
  for (int k = 0; k < 12; k++) {
    N->woff[k] = off;
    off += 27 * N->cin(k) * kLayers[k].cout;
    N->boff[k] = off;
    off += kLayers[k].cout;
  }
  N->nparam = off;
  // k-major copies for the layers the MFMA convolution takes (cin % 8 == 0,
  // cout a multiple of 16)
  int64_t wto = 0;
  for (int k = 0; k < 12; k++) {
    const int cin = int(N->cin(k)), cout = kLayers[k].cout;
    if (cin % 8 == 0 && kLayers[k].cin_a % 8 == 0 && cout % 16 == 0) {
      N->wtoff[k] = wto;
      wto += int64_t(cout) * sxu::mfma_kpad(cin);
    } else {
      N->wtoff[k] = -1;
    }
  }
  N->Wt.alloc(size_t(wto) * 2);
  N->P.alloc(size_t(off) * 4);
  N->G.alloc(size_t(off) * 4);
  N->M.alloc(size_t(off) * 4);
  N->V.alloc(size_t(off) * 4);
  N->Wbf.alloc(size_t(off) * 2);
  init_params(N, seed);
  for (int k = 0; k < 12; k++) {
    const int64_t n = N->nvox[kLayers[k].lout] * kLayers[k].cout;
    if (k < 11) N->act[k].alloc(size_t(n) * 2);
    N->gact[k].alloc(size_t(n) * 4);
  }
  for (int j = 0; j < 3; j++) {
    const int64_t n = N->nvox[kLayers[8 + j].lout] * kLayers[8 + j].cin_a;
    N->up[j].alloc(size_t(n) * 2);
    N->gup[j].alloc(size_t(n) * 4);
  }
  N->xin.alloc(size_t(N->nvox[0]) * 3 * 2);
  N->yout.alloc(size_t(N->nvox[0]) * 3 * 4);
  int64_t wmax = 0, tmax = 0;
  for (int k = 0; k < 12; k++) {
    wmax = std::max<int64_t>(wmax, 27 * N->cin(k) * kLayers[k].cout);
    tmax = std::max<int64_t>(tmax, N->nvox[kLayers[k].lin] * kLayers[k].cin_a);
  }
  N->wflip.alloc(size_t(wmax) * 2);
  N->tmp.alloc(size_t(tmax) * 4);
  N->sums.alloc(2 * sizeof(double));
  UN_HIP(hipStreamCreateWithFlags(&N->stream, hipStreamNonBlocking));
}

// loss of the current output against target t (device FP32 [n0][3]); writes
// dL/dy into gact[11]
double loss_and_grad(sx_unet* N, const float* t)
{
  UN_HIP(hipMemsetAsync(N->sums.p, 0, 2 * sizeof(double), N->stream));
  UN_HIP(sxu::loss(N->yout.as<float>(), t, N->nvox[0] * 3, N->sums.as<double>(), N->gact[11].as<float>(), N->stream));
  double h[2];
  UN_HIP(hipMemcpyAsync(h, N->sums.p, sizeof(h), hipMemcpyDeviceToHost, N->stream));
  UN_HIP(hipStreamSynchronize(N->stream));
  return h[1] > 0 ? h[0] / h[1] : 0.0;
}

}  // namespace

extern "C" {

int sx_unet_create(int nx, int ny, int nz, uint64_t seed, sx_unet** out)
{
  return unet_api([&] {
    if (!out) throw unet_failure(SX_ERR_ARG, "null output");
    auto N = std::make_unique<sx_unet>();
    init(N.get(), nx, ny, nz, seed);
    *out = N.release();
  });
}

int sx_unet_set_activation(sx_unet* N, int activation, uint64_t seed)
{
  return unet_api([&] {
    if (!N) throw unet_failure(SX_ERR_ARG, "null network");
    if (activation != SX_UNET_RELU && activation != SX_UNET_IDENTITY)
      throw unet_failure(SX_ERR_ARG, "activation is SX_UNET_RELU or SX_UNET_IDENTITY");
    N->identity = activation == SX_UNET_IDENTITY ? 1 : 0;
    init_params(N, seed);
  });
}

int sx_unet_destroy(sx_unet* N)
{
  if (N && N->stream) (void)hipStreamDestroy(N->stream);
  delete N;
  return SX_OK;
}

int sx_unet_info(const sx_unet* N, int64_t* nparam, int64_t* nvox0)
{
  return unet_api([&] {
    if (!N) throw unet_failure(SX_ERR_ARG, "null network");
    if (nparam) *nparam = N->nparam;
    if (nvox0) *nvox0 = N->nvox[0];
  });
}

int sx_unet_get_params(const sx_unet* N, float* host)
{
  return unet_api([&] {
    UN_HIP(hipStreamSynchronize(N->stream));
    UN_HIP(hipMemcpy(host, N->P.p, size_t(N->nparam) * 4, hipMemcpyDeviceToHost));
  });
}

int sx_unet_set_params(sx_unet* N, const float* host)
{
  return unet_api([&] {
    UN_HIP(hipStreamSynchronize(N->stream));
    UN_HIP(hipMemcpy(N->P.p, host, size_t(N->nparam) * 4, hipMemcpyHostToDevice));
    UN_HIP(sxu::to_bf16(N->P.as<float>(), N->Wbf.as<uint16_t>(), N->nparam, N->stream));
    UN_HIP(hipStreamSynchronize(N->stream));
    N->wt_dirty = true;
  });
}

int sx_unet_get_grads(const sx_unet* N, float* host)
{
  return unet_api([&] {
    UN_HIP(hipStreamSynchronize(N->stream));
    UN_HIP(hipMemcpy(host, N->G.p, size_t(N->nparam) * 4, hipMemcpyDeviceToHost));
  });
}

// y = N(x): x, y device FP32 [n0][3] (x rounded to BF16 on entry)
int sx_unet_forward(sx_unet* N, const float* x, float* y)
{
  return unet_api([&] {
    if (!N || !x || !y) throw unet_failure(SX_ERR_ARG, "null argument");
    UN_HIP(sxu::to_bf16(x, N->xin.as<uint16_t>(), N->nvox[0] * 3, N->stream));
    forward(N);
    UN_HIP(hipMemcpyAsync(y, N->yout.p, size_t(N->nvox[0]) * 3 * 4, hipMemcpyDeviceToDevice, N->stream));
    UN_HIP(hipStreamSynchronize(N->stream));
  });
}

// one forward + backward on (x, target): loss = |N(x) - t|^2 / |t|^2; the
// parameter gradients accumulate (sx_unet_adam applies and clears them)
int sx_unet_backward(sx_unet* N, const float* x, const float* t, double* loss)
{
  return unet_api([&] {
    if (!N || !x || !t) throw unet_failure(SX_ERR_ARG, "null argument");
    UN_HIP(sxu::to_bf16(x, N->xin.as<uint16_t>(), N->nvox[0] * 3, N->stream));
    forward(N);
    const double l = loss_and_grad(N, t);
    backward(N);
    UN_HIP(hipStreamSynchronize(N->stream));
    if (loss) *loss = l;
  });
}

int sx_unet_adam(sx_unet* N, double lr)
{
  return unet_api([&] {
    if (!N) throw unet_failure(SX_ERR_ARG, "null network");
    N->step++;
    // biases have no BF16 copy in use, but share the offsets; refresh all
    UN_HIP(sxu::adam(N->P.as<float>(), N->M.as<float>(), N->V.as<float>(), N->G.as<float>(), N->Wbf.as<uint16_t>(),
                     N->nparam, float(lr), 0.9f, 0.999f, 1e-8f, N->step, N->stream));
    UN_HIP(hipStreamSynchronize(N->stream));
    N->wt_dirty = true;
  });
}

int sx_unet_set_mfma(sx_unet* N, int on)
{
  return unet_api([&] {
    if (!N) throw unet_failure(SX_ERR_ARG, "null network");
    N->use_mfma = on ? 1 : 0;
    N->wt_dirty = true;
  });
}

}  // extern "C"
