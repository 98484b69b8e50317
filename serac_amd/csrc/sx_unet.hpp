// sx_unet.hpp -- kernels of the learned coarse-grid preconditioner
// (sx_unet.hip) and the network that drives them (sx_learn.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sxu {

struct Dims3 {
  int n[3];
};

// order-p nodal grid of a structured mesh (nn = p nv - p + 1 nodes per
// direction) over its vertex grid nv; xi: the p + 1 GLL points on [0, 1]
struct FEGrid {
  int    p;
  int    nv[3], nn[3];
  double xi[6];
};

// one 3x3x3 convolution; x = concat(xa, xb) along channels, channels-last
struct ConvArgs {
  const void*     xa;
  int             ca;
  const void*     xb;  // may be null (cb = 0)
  int             cb;
  int             din[3], dout[3];
  int             mode;  // 0: stride 1; 1: stride 2; 2: transpose of stride 2 (weights flipped)
  const uint16_t* w;     // BF16 [27][ca + cb][cout]
  const float*    bias;  // [cout] or null
  int             relu;
  void*           y;  // [nout][cout]
};

struct WGradArgs {
  const void*  xa;
  int          ca;
  const void*  xb;
  int          cb;
  int          din[3], dout[3];
  int          mode;  // of the forward convolution (0 or 1)
  const float* g;     // dL/d(pre-activation) [nout][cout]
  float*       dw;    // [27][ca + cb][cout], accumulated into
  float*       db;    // [cout], accumulated into (may be null)
};

hipError_t conv(const ConvArgs& A, int cout, bool in_f32, bool out_f32, hipStream_t s);
// the matrix-core (MFMA) implicit-GEMM form of conv for BF16 input and output,
// modes 0 and 1, cin % 8 == 0 and cout in {16, 32, 64}; wt holds the weights
// k-contiguous per output channel ([cout][mfma_kpad(cin)], weights_kmajor)
int        mfma_kpad(int cin);
bool       conv_mfma_supported(const ConvArgs& A, int cout, bool in_f32, bool out_f32);
hipError_t weights_kmajor(const uint16_t* w, int cin, int cout, uint16_t* wt, hipStream_t s);
hipError_t conv_mfma(const ConvArgs& A, int cout, const uint16_t* wt, hipStream_t s);
hipError_t conv_wgrad(const WGradArgs& A, int cout, bool in_f32, hipStream_t s);
hipError_t flip_weights(const uint16_t* w, int cin, int cout, int c0, int cpart, uint16_t* wt, hipStream_t s);
hipError_t relu_bwd(float* g, const uint16_t* y, int64_t n, hipStream_t s);
hipError_t prolong_bf16(const uint16_t* xc, Dims3 dc, uint16_t* yf, Dims3 df, int C, hipStream_t s);
hipError_t restrict_f32(const float* gf, Dims3 df, float* gc, Dims3 dc, int C, hipStream_t s);
hipError_t fill_f32(float* x, float v, int64_t n, hipStream_t s);
hipError_t add_f32(float* y, const float* x, int64_t n, hipStream_t s);
hipError_t loss(const float* y, const float* t, int64_t n, double* sums, float* dy, hipStream_t s);
hipError_t adam(float* w, float* m, float* v, float* g, uint16_t* wbf, int64_t n, float lr, float b1, float b2,
                float eps, int step, hipStream_t s);
hipError_t to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t s);
// rc = P^T r (or, with `average`, (P^T r) / (P^T 1) row by row);
// z += gain sqrt(ss / nss) P ec
hipError_t fe_restrict(const FEGrid& G, const double* r, float* rc, hipStream_t s, bool average = false);
hipError_t fe_prolong_add(const FEGrid& G, const float* ec, const double* ss, int64_t nss, double* z, hipStream_t s,
                          double gain = 1.0);
// ss[0] += sum x^2;  y = x / sqrt(ss / nss)
hipError_t sumsq(const float* x, int64_t n, double* ss, hipStream_t s);
hipError_t normalize(const float* x, const double* ss, int64_t nss, int64_t n, float* y, hipStream_t s);
// y = a - b formed in FP64, stored as FP32
hipError_t sub_f64(const double* a, const double* b, float* y, int64_t n, hipStream_t s);
// xc = (P^T x) / (P^T 1) in FP64
hipError_t fe_average(const FEGrid& G, const double* x, double* xc, hipStream_t s);

}  // namespace sxu
