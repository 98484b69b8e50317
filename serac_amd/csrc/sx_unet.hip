// sx_unet.hip -- the learned coarse-grid preconditioner: a 4-level 3D
// convolutional U-Net (16 -> 32 -> 64 -> 64 channels, 3x3x3 kernels, BF16
// weights and activations, FP32 accumulation) on the vertex grid of a
// structured hex mesh, with its training kernels (backward, Adam) and the
// coupling to the order-p nodal residual.
//
// Tensors are channels-last, [z][y][x][c], on vertex-centred grids: level
// l + 1 has (n_l - 1) / 2 + 1 nodes per direction and its node i sits on node
// 2i of level l.  Every convolution is 3x3x3 with zero padding; the
// encoder's down convolutions have stride 2, the decoder prolongs (trilinear,
// vertex-centred) and convolves the concatenation [prolonged, skip].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "sx_unet.hpp"

namespace sxu {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f)
{
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float ld(const uint16_t* p, int64_t i) { return bf2f(p[i]); }
__device__ __forceinline__ float ld(const float* p, int64_t i) { return p[i]; }
__device__ __forceinline__ void st(uint16_t* p, int64_t i, float v) { p[i] = f2bf(v); }
__device__ __forceinline__ void st(float* p, int64_t i, float v) { p[i] = v; }

inline unsigned blocks_for(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

// source node of output node o along one direction for tap offset d in {-1,0,1}
//   mode 0: o + d; mode 1: 2o + d; mode 2: (o + d) / 2 when o + d is even
__device__ __forceinline__ int src_index(int o, int d, int mode, int n_in)
{
  int s;
  if (mode == 0)
    s = o + d;
  else if (mode == 1)
    s = 2 * o + d;
  else {
    const int t = o + d;
    if (t & 1) return -1;
    s = t / 2;
  }
  return (s >= 0 && s < n_in) ? s : -1;
}

// y[o][co] = bias[co] + sum_{tap, ci} w[tap][ci][co] x[src(o, tap)][ci],
// x = [xa (ca channels), xb (cb channels)]; optional ReLU.  One thread per
// output node; the tap's weights are staged in LDS.
template <int COUT, typename TI, typename TO>
__global__ void __launch_bounds__(kThreads) k_conv(ConvArgs A)
{
  extern __shared__ uint16_t wl[];  // [ca + cb][COUT] of the current tap
  const int     cin = A.ca + A.cb;
  const int64_t nout = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const int64_t o    = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const bool    act  = o < nout;
  int           ox = 0, oy = 0, oz = 0;
  if (act) {
    ox = int(o % A.dout[0]);
    oy = int((o / A.dout[0]) % A.dout[1]);
    oz = int(o / (int64_t(A.dout[0]) * A.dout[1]));
  }
  float acc[COUT];
#pragma unroll
  for (int c = 0; c < COUT; c++) acc[c] = 0.0f;
  const uint16_t* W  = static_cast<const uint16_t*>(A.w);
  const TI*       xa = static_cast<const TI*>(A.xa);
  const TI*       xb = static_cast<const TI*>(A.xb);
  for (int tap = 0; tap < 27; tap++) {
    __syncthreads();
    for (int t = threadIdx.x; t < cin * COUT; t += kThreads) wl[t] = W[int64_t(tap) * cin * COUT + t];
    __syncthreads();
    if (!act) continue;
    const int sx = src_index(ox, tap % 3 - 1, A.mode, A.din[0]);
    const int sy = src_index(oy, (tap / 3) % 3 - 1, A.mode, A.din[1]);
    const int sz = src_index(oz, tap / 9 - 1, A.mode, A.din[2]);
    if (sx < 0 || sy < 0 || sz < 0) continue;
    const int64_t s = (int64_t(sz) * A.din[1] + sy) * A.din[0] + sx;
    for (int ci = 0; ci < A.ca; ci++) {
      const float     v  = ld(xa, s * A.ca + ci);
      const uint16_t* wr = wl + ci * COUT;
#pragma unroll
      for (int c = 0; c < COUT; c++) acc[c] += v * bf2f(wr[c]);
    }
    for (int ci = 0; ci < A.cb; ci++) {
      const float     v  = ld(xb, s * A.cb + ci);
      const uint16_t* wr = wl + (A.ca + ci) * COUT;
#pragma unroll
      for (int c = 0; c < COUT; c++) acc[c] += v * bf2f(wr[c]);
    }
  }
  if (!act) return;
  TO* y = static_cast<TO*>(A.y);
#pragma unroll
  for (int c = 0; c < COUT; c++) {
    float v = acc[c] + (A.bias ? A.bias[c] : 0.0f);
    if (A.relu) v = v > 0.0f ? v : 0.0f;
    st(y, o * COUT + c, v);
  }
}

// The same convolution with one lane per (output node, output channel):
// 256 / COUT nodes per workgroup, the tap's weights and the nodes' input
// values staged in LDS.  Used on the coarse levels, where one thread per node
// would leave most of the GPU idle.
template <int COUT, typename TI, typename TO>
__global__ void __launch_bounds__(kThreads) k_conv_cp(ConvArgs A)
{
  constexpr int VPB = kThreads / COUT;
  extern __shared__ float cp_smem[];  // xs: [VPB][cin] FP32, then wl: [cin][COUT] BF16
  const int      cin  = A.ca + A.cb;
  float*         xs   = cp_smem;
  uint16_t*      wl   = reinterpret_cast<uint16_t*>(cp_smem + VPB * cin);
  const int64_t  nout = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const int      c    = int(threadIdx.x) % COUT;
  const int      vl   = int(threadIdx.x) / COUT;
  const int64_t  o    = int64_t(blockIdx.x) * VPB + vl;
  const uint16_t* W   = static_cast<const uint16_t*>(A.w);
  const TI*      xa   = static_cast<const TI*>(A.xa);
  const TI*      xb   = static_cast<const TI*>(A.xb);
  float          acc  = 0.0f;
  for (int tap = 0; tap < 27; tap++) {
    __syncthreads();
    for (int t = threadIdx.x; t < cin * COUT; t += kThreads) wl[t] = W[int64_t(tap) * cin * COUT + t];
    for (int t = threadIdx.x; t < VPB * cin; t += kThreads) {
      const int     v  = t / cin, ci = t - v * cin;
      const int64_t oo = int64_t(blockIdx.x) * VPB + v;
      float         xv = 0.0f;
      if (oo < nout) {
        const int ox = int(oo % A.dout[0]);
        const int oy = int((oo / A.dout[0]) % A.dout[1]);
        const int oz = int(oo / (int64_t(A.dout[0]) * A.dout[1]));
        const int sx = src_index(ox, tap % 3 - 1, A.mode, A.din[0]);
        const int sy = src_index(oy, (tap / 3) % 3 - 1, A.mode, A.din[1]);
        const int sz = src_index(oz, tap / 9 - 1, A.mode, A.din[2]);
        if (sx >= 0 && sy >= 0 && sz >= 0) {
          const int64_t sidx = (int64_t(sz) * A.din[1] + sy) * A.din[0] + sx;
          xv = ci < A.ca ? ld(xa, sidx * A.ca + ci) : ld(xb, sidx * A.cb + (ci - A.ca));
        }
      }
      xs[t] = xv;
    }
    __syncthreads();
    const float* xv = xs + vl * cin;
    for (int ci = 0; ci < cin; ci++) acc += xv[ci] * bf2f(wl[ci * COUT + c]);
  }
  if (o >= nout) return;
  float v = acc + (A.bias ? A.bias[c] : 0.0f);
  if (A.relu) v = v > 0.0f ? v : 0.0f;
  st(static_cast<TO*>(A.y), o * COUT + c, v);
}

// The same convolution as an implicit GEMM on the matrix cores: rows are
// output nodes, columns output channels, K = (tap, input channel) in 27 * cin,
// padded to a multiple of 32.  mfma_f32_16x16x32_bf16 tiles: lane l holds
// A(row l % 16, k 8 (l / 16) .. + 7) -- eight consecutive channels of one tap of
// one node, a single 16-byte load from the channels-last input -- and
// B(k 8 (l / 16) .. + 7, column l % 16) from the weights stored k-contiguous
// per output channel (wt[co][Kpad]); D(row 4 (l / 16) + r, column l % 16).
// Each wave computes RT row tiles (16 RT nodes) x COUT / 16 column tiles; FP32
// accumulation, bias and ReLU in the epilogue, BF16 output.  Requires BF16
// input and output, stride 1 or 2 (modes 0, 1) and cin % 8 == 0.
#ifndef MFMA_RT16
#define MFMA_RT16 1
#endif
#ifndef MFMA_RT32
#define MFMA_RT32 1
#endif
#ifndef MFMA_RT64
#define MFMA_RT64 1
#endif
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float  f32x4_t __attribute__((ext_vector_type(4)));

template <int COUT, int RT>
__global__ void __launch_bounds__(kThreads) k_conv_mfma(ConvArgs A, const uint16_t* wt, int kpad)
{
  constexpr int NT = COUT / 16;
  const int     cin  = A.ca + A.cb;
  const int64_t nout = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const int     lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
  const int     m = lane & 15, kg = lane >> 4;
  const int64_t base = (int64_t(blockIdx.x) * (kThreads / 64) + wave) * (RT * 16);
  int           ox[RT], oy[RT], oz[RT];
  bool          ok[RT];
#pragma unroll
  for (int rt = 0; rt < RT; rt++) {
    const int64_t o = base + rt * 16 + m;
    ok[rt]          = o < nout;
    const int64_t oc = ok[rt] ? o : 0;
    ox[rt]           = int(oc % A.dout[0]);
    oy[rt]           = int((oc / A.dout[0]) % A.dout[1]);
    oz[rt]           = int(oc / (int64_t(A.dout[0]) * A.dout[1]));
  }
  f32x4_t acc[RT][NT];
#pragma unroll
  for (int rt = 0; rt < RT; rt++)
#pragma unroll
    for (int ct = 0; ct < NT; ct++) acc[rt][ct] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  const uint16_t* xa = static_cast<const uint16_t*>(A.xa);
  const uint16_t* xb = static_cast<const uint16_t*>(A.xb);
  const int       K  = 27 * cin;
  for (int k0 = 8 * kg; k0 < kpad; k0 += 32) {
    bf16x8_t av[RT];
    const int tap = k0 / cin, ci = k0 - tap * cin;
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
      av[rt] = bf16x8_t{};
      if (k0 < K && ok[rt]) {
        const int sx = src_index(ox[rt], tap % 3 - 1, A.mode, A.din[0]);
        const int sy = src_index(oy[rt], (tap / 3) % 3 - 1, A.mode, A.din[1]);
        const int sz = src_index(oz[rt], tap / 9 - 1, A.mode, A.din[2]);
        if (sx >= 0 && sy >= 0 && sz >= 0) {
          const int64_t   sidx = (int64_t(sz) * A.din[1] + sy) * A.din[0] + sx;
          const uint16_t* src  = ci < A.ca ? xa + sidx * A.ca + ci : xb + sidx * A.cb + (ci - A.ca);
          av[rt]               = *reinterpret_cast<const bf16x8_t*>(src);
        }
      }
    }
#pragma unroll
    for (int ct = 0; ct < NT; ct++) {
      const bf16x8_t bv = *reinterpret_cast<const bf16x8_t*>(wt + int64_t(ct * 16 + m) * kpad + k0);
#pragma unroll
      for (int rt = 0; rt < RT; rt++) acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rt], bv, acc[rt][ct], 0, 0, 0);
    }
  }
  uint16_t* y = static_cast<uint16_t*>(A.y);
#pragma unroll
  for (int rt = 0; rt < RT; rt++)
#pragma unroll
    for (int ct = 0; ct < NT; ct++) {
      const int   co = ct * 16 + m;
      const float b  = A.bias ? A.bias[co] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t o = base + rt * 16 + 4 * kg + r;
        if (o >= nout) continue;
        float v = acc[rt][ct][r] + b;
        if (A.relu) v = v > 0.0f ? v : 0.0f;
        y[o * COUT + co] = f2bf(v);
      }
    }
}

// The same with one 16 x 16 output tile (16 nodes, 16 channels) per wave, for
// levels with few nodes: waves = row tiles x COUT / 16.
template <int COUT>
__global__ void __launch_bounds__(kThreads) k_conv_mfma_tile(ConvArgs A, const uint16_t* wt, int kpad)
{
  constexpr int NT   = COUT / 16;
  const int     cin  = A.ca + A.cb;
  const int64_t nout = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const int     lane = int(threadIdx.x) & 63;
  const int     m = lane & 15, kg = lane >> 4;
  const int64_t w    = int64_t(blockIdx.x) * (kThreads / 64) + (int(threadIdx.x) >> 6);
  const int64_t nrow = (nout + 15) / 16;
  if (w >= nrow * NT) return;  // whole wave: no barrier follows
  const int64_t base = (w / NT) * 16;
  const int     ct   = int(w % NT);
  const int64_t o0   = base + m;
  const bool    ok   = o0 < nout;
  const int64_t oc   = ok ? o0 : 0;
  const int     ox = int(oc % A.dout[0]), oy = int((oc / A.dout[0]) % A.dout[1]);
  const int     oz = int(oc / (int64_t(A.dout[0]) * A.dout[1]));
  f32x4_t       acc = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  const uint16_t* xa = static_cast<const uint16_t*>(A.xa);
  const uint16_t* xb = static_cast<const uint16_t*>(A.xb);
  const uint16_t* wr = wt + int64_t(ct * 16 + m) * kpad;
  const int       K  = 27 * cin;
  for (int k0 = 8 * kg; k0 < kpad; k0 += 32) {
    const int tap = k0 / cin, ci = k0 - tap * cin;
    bf16x8_t  av  = bf16x8_t{};
    if (k0 < K && ok) {
      const int sx = src_index(ox, tap % 3 - 1, A.mode, A.din[0]);
      const int sy = src_index(oy, (tap / 3) % 3 - 1, A.mode, A.din[1]);
      const int sz = src_index(oz, tap / 9 - 1, A.mode, A.din[2]);
      if (sx >= 0 && sy >= 0 && sz >= 0) {
        const int64_t   sidx = (int64_t(sz) * A.din[1] + sy) * A.din[0] + sx;
        const uint16_t* src  = ci < A.ca ? xa + sidx * A.ca + ci : xb + sidx * A.cb + (ci - A.ca);
        av                   = *reinterpret_cast<const bf16x8_t*>(src);
      }
    }
    const bf16x8_t bv = *reinterpret_cast<const bf16x8_t*>(wr + k0);
    acc               = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  }
  uint16_t*   y  = static_cast<uint16_t*>(A.y);
  const int   co = ct * 16 + m;
  const float bb = A.bias ? A.bias[co] : 0.0f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int64_t o = base + 4 * kg + r;
    if (o >= nout) continue;
    float v = acc[r] + bb;
    if (A.relu) v = v > 0.0f ? v : 0.0f;
    y[o * COUT + co] = f2bf(v);
  }
}

// wt[co][k] = w[k / cin][k % cin][co] for k < 27 cin, 0 up to kpad
__global__ void k_weights_kmajor(const uint16_t* w, int cin, int cout, int kpad, uint16_t* wt)
{
  const int64_t n = int64_t(cout) * kpad;
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kThreads) {
    const int co = int(t / kpad), k = int(t % kpad);
    wt[t]        = k < 27 * cin ? w[(int64_t(k / cin) * cin + k % cin) * cout + co] : uint16_t(0);
  }
}

// dw[tap][ci][co] += sum_o x[src(o, tap)][ci] g[o][co] (x = [xa, xb]) and, for
// tap 0, db[co] += sum_o g[o][co].  grid = (ceil(nout / (64 chunks)), 27): each
// workgroup sums `chunks` runs of 64 output nodes in registers before one
// atomic add per weight.
template <int COUT, typename TI>
__global__ void __launch_bounds__(kThreads) k_conv_wgrad(WGradArgs A, int chunks)
{
  constexpr int VC   = 64;
  constexpr int EMAX = (128 * COUT + kThreads - 1) / kThreads;  // weights per thread at cin <= 128
  __shared__ float sg[VC * COUT];
  __shared__ float sx[VC * 128];
  const int     cin  = A.ca + A.cb;
  const int     tap  = blockIdx.y;
  const int64_t nout = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const TI*     xa   = static_cast<const TI*>(A.xa);
  const TI*     xb   = static_cast<const TI*>(A.xb);
  float         part[EMAX];
#pragma unroll
  for (int k = 0; k < EMAX; k++) part[k] = 0.0f;
  float bpart = 0.0f;
  for (int ch = 0; ch < chunks; ch++) {
    const int64_t o0 = (int64_t(blockIdx.x) * chunks + ch) * VC;
    if (o0 >= nout) break;
    __syncthreads();
    for (int t = threadIdx.x; t < VC * COUT; t += kThreads) {
      const int64_t o = o0 + t / COUT;
      sg[t]           = o < nout ? A.g[o * COUT + t % COUT] : 0.0f;
    }
    for (int j = threadIdx.x; j < VC; j += kThreads) {
      const int64_t o = o0 + j;
      int64_t       s = -1;
      if (o < nout) {
        const int ox = int(o % A.dout[0]), oy = int((o / A.dout[0]) % A.dout[1]);
        const int oz = int(o / (int64_t(A.dout[0]) * A.dout[1]));
        const int ix = src_index(ox, tap % 3 - 1, A.mode, A.din[0]);
        const int iy = src_index(oy, (tap / 3) % 3 - 1, A.mode, A.din[1]);
        const int iz = src_index(oz, tap / 9 - 1, A.mode, A.din[2]);
        if (ix >= 0 && iy >= 0 && iz >= 0) s = (int64_t(iz) * A.din[1] + iy) * A.din[0] + ix;
      }
      for (int ci = 0; ci < A.ca; ci++) sx[j * 128 + ci] = s >= 0 ? ld(xa, s * A.ca + ci) : 0.0f;
      for (int ci = 0; ci < A.cb; ci++) sx[j * 128 + A.ca + ci] = s >= 0 ? ld(xb, s * A.cb + ci) : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < EMAX; k++) {
      const int e = int(threadIdx.x) + k * kThreads;
      if (e < cin * COUT) {
        const int ci = e / COUT, co = e % COUT;
        float     s  = 0.0f;
        for (int j = 0; j < VC; j++) s += sx[j * 128 + ci] * sg[j * COUT + co];
        part[k] += s;
      }
    }
    if (tap == 0 && int(threadIdx.x) < COUT)
      for (int j = 0; j < VC; j++) bpart += sg[j * COUT + threadIdx.x];
  }
#pragma unroll
  for (int k = 0; k < EMAX; k++) {
    const int e = int(threadIdx.x) + k * kThreads;
    if (e < cin * COUT) atomicAdd(&A.dw[int64_t(tap) * cin * COUT + e], part[k]);
  }
  if (tap == 0 && A.db && int(threadIdx.x) < COUT) atomicAdd(&A.db[threadIdx.x], bpart);
}

// wt[tap'][co][ci] = w[26 - tap'][c0 + ci][co] for ci < cpart: the weights of
// the data-gradient convolution of input part [c0, c0 + cpart)
__global__ void k_flip_weights(const uint16_t* w, int cin, int cout, int c0, int cpart, uint16_t* wt)
{
  const int64_t n = int64_t(27) * cout * cpart;
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kThreads) {
    const int ci = int(t % cpart), co = int((t / cpart) % cout), tp = int(t / (int64_t(cpart) * cout));
    wt[t]        = w[(int64_t(26 - tp) * cin + c0 + ci) * cout + co];
  }
}

// g[i] *= (y[i] > 0): the ReLU derivative from the activation's output
__global__ void k_relu_bwd(float* g, const uint16_t* y, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    if (!(bf2f(y[i]) > 0.0f)) g[i] = 0.0f;
}

// vertex-centred trilinear prolongation, coarse nc^3 -> fine nf^3, C channels:
// fine node f takes coarse f/2 when f is even and the mean of (f-1)/2 and
// (f+1)/2 when odd, per direction
template <typename TI, typename TO>
__global__ void k_prolong(const TI* xc, Dims3 dc, TO* yf, Dims3 df, int C, int accumulate)
{
  const int64_t n = int64_t(df.n[0]) * df.n[1] * df.n[2] * C;
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kThreads) {
    const int     c = int(t % C);
    const int64_t f = t / C;
    const int     fx = int(f % df.n[0]), fy = int((f / df.n[0]) % df.n[1]), fz = int(f / (int64_t(df.n[0]) * df.n[1]));
    int           lo[3], hi[3];
    const int     fv[3] = {fx, fy, fz};
    float         w[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = fv[d] / 2;
      hi[d] = (fv[d] & 1) ? lo[d] + 1 : lo[d];
      w[d]  = (fv[d] & 1) ? 0.5f : 1.0f;
    }
    float s = 0.0f;
    for (int k = 0; k < 8; k++) {
      const int bx = k & 1, by = (k >> 1) & 1, bz = k >> 2;
      if ((bx && hi[0] == lo[0]) || (by && hi[1] == lo[1]) || (bz && hi[2] == lo[2])) continue;
      const int   cx = bx ? hi[0] : lo[0], cy = by ? hi[1] : lo[1], cz = bz ? hi[2] : lo[2];
      const float ww = w[0] * w[1] * w[2];
      s += ww * ld(xc, ((int64_t(cz) * dc.n[1] + cy) * dc.n[0] + cx) * C + c);
    }
    if (accumulate)
      st(yf, t, ld(yf, t) + s);
    else
      st(yf, t, s);
  }
}

// transpose of k_prolong: coarse node o gathers fine nodes 2o + d, d in
// {-1, 0, 1}, with weights {0.5, 1, 0.5} per direction
__global__ void k_restrict(const float* gf, Dims3 df, float* gc, Dims3 dc, int C)
{
  const int64_t n = int64_t(dc.n[0]) * dc.n[1] * dc.n[2] * C;
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kThreads) {
    const int     c = int(t % C);
    const int64_t o = t / C;
    const int     ov[3] = {int(o % dc.n[0]), int((o / dc.n[0]) % dc.n[1]), int(o / (int64_t(dc.n[0]) * dc.n[1]))};
    float         s     = 0.0f;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int fx = 2 * ov[0] + dx, fy = 2 * ov[1] + dy, fz = 2 * ov[2] + dz;
          if (fx < 0 || fy < 0 || fz < 0 || fx >= df.n[0] || fy >= df.n[1] || fz >= df.n[2]) continue;
          const float w = (dx ? 0.5f : 1.0f) * (dy ? 0.5f : 1.0f) * (dz ? 0.5f : 1.0f);
          s += w * gf[((int64_t(fz) * df.n[1] + fy) * df.n[0] + fx) * C + c];
        }
    gc[t] = s;
  }
}

__global__ void k_fill_f(float* x, float v, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) x[i] = v;
}
__global__ void k_add_f(float* y, const float* x, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) y[i] += x[i];
}

// relative squared error: sums[0] += (y - t)^2, sums[1] += t^2 (FP64)
__global__ void k_loss_sums(const float* y, const float* t, int64_t n, double* sums)
{
  __shared__ double s0[kThreads], s1[kThreads];
  double a = 0.0, b = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const double d = double(y[i]) - double(t[i]);
    a += d * d;
    b += double(t[i]) * double(t[i]);
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&sums[0], s0[0]);
    atomicAdd(&sums[1], s1[0]);
  }
}
// dy = 2 (y - t) / sum t^2
__global__ void k_loss_grad(const float* y, const float* t, int64_t n, const double* sums, float* dy)
{
  const double inv = sums[1] > 0 ? 2.0 / sums[1] : 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    dy[i] = float((double(y[i]) - double(t[i])) * inv);
}

// Adam on FP32 master weights; refreshes the BF16 copy and clears the gradient
__global__ void k_adam(float* w, float* m, float* v, float* g, uint16_t* wbf, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i]           = mi;
    v[i]           = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i]           = wi;
    if (wbf) wbf[i] = f2bf(wi);
    g[i] = 0.0f;
  }
}

__global__ void k_to_bf16(const float* x, uint16_t* y, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    y[i] = f2bf(x[i]);
}

// ---- coupling to the order-p nodal space of a structured mesh ----
// P interpolates vertex (Q1) values to the order-p nodes; in 1D, node j and
// vertex i couple through w(j - p i), w(d) = 1 - xi_d for 0 <= d <= p and
// xi_{p+d} for -p <= d < 0 (xi: the GLL points on [0, 1]); P is the tensor
// product.  Fine vectors are byVDIM FP64 with 3 components.
__device__ __forceinline__ double fe_w(const FEGrid& G, int d)
{
  return d >= 0 ? 1.0 - G.xi[d] : G.xi[G.p + d];
}

// rc = P^T r (FP32 out); with `average`, each row is divided by its weight
// sum (P^T 1), the hat-weighted mean of the nodal values around the vertex
template <typename TO>
__global__ void k_fe_restrict(FEGrid G, const double* r, TO* rc, int average)
{
  const int64_t nv = int64_t(G.nv[0]) * G.nv[1] * G.nv[2];
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nv * 3; t += int64_t(gridDim.x) * kThreads) {
    const int     c = int(t % 3);
    const int64_t i = t / 3;
    const int     iv[3] = {int(i % G.nv[0]), int((i / G.nv[0]) % G.nv[1]), int(i / (int64_t(G.nv[0]) * G.nv[1]))};
    double        s = 0.0, ws = 0.0;
    for (int dz = -G.p; dz <= G.p; dz++) {
      const int jz = G.p * iv[2] + dz;
      if (jz < 0 || jz >= G.nn[2]) continue;
      const double wz = fe_w(G, dz);
      for (int dy = -G.p; dy <= G.p; dy++) {
        const int jy = G.p * iv[1] + dy;
        if (jy < 0 || jy >= G.nn[1]) continue;
        const double wzy = wz * fe_w(G, dy);
        for (int dx = -G.p; dx <= G.p; dx++) {
          const int jx = G.p * iv[0] + dx;
          if (jx < 0 || jx >= G.nn[0]) continue;
          const double w = wzy * fe_w(G, dx);
          s += w * r[((int64_t(jz) * G.nn[1] + jy) * G.nn[0] + jx) * 3 + c];
          ws += w;
        }
      }
    }
    rc[t] = TO(average ? s / ws : s);
  }
}

// z += sqrt(ss / n) * P ec  (ss: the sum of squares the input was normalized by)
__global__ void k_fe_prolong_add(FEGrid G, const float* ec, const double* ss, int64_t nss, double gain, double* z)
{
  const double  scale = ss[0] > 0 ? gain * sqrt(ss[0] / double(nss)) : 0.0;
  const int64_t nn    = int64_t(G.nn[0]) * G.nn[1] * G.nn[2];
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nn * 3; t += int64_t(gridDim.x) * kThreads) {
    const int     c = int(t % 3);
    const int64_t j = t / 3;
    const int     jv[3] = {int(j % G.nn[0]), int((j / G.nn[0]) % G.nn[1]), int(j / (int64_t(G.nn[0]) * G.nn[1]))};
    int           e[3], k[3];
    for (int d = 0; d < 3; d++) {
      e[d] = jv[d] / G.p;
      k[d] = jv[d] - G.p * e[d];
    }
    double s = 0.0;
    for (int b = 0; b < 8; b++) {
      const int bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
      if ((bx && k[0] == 0) || (by && k[1] == 0) || (bz && k[2] == 0)) continue;
      const double w = (bx ? G.xi[k[0]] : 1.0 - G.xi[k[0]]) * (by ? G.xi[k[1]] : 1.0 - G.xi[k[1]]) *
                       (bz ? G.xi[k[2]] : 1.0 - G.xi[k[2]]);
      const int64_t v = (int64_t(e[2] + bz) * G.nv[1] + (e[1] + by)) * G.nv[0] + (e[0] + bx);
      s += w * double(ec[v * 3 + c]);
    }
    z[t] += scale * s;
  }
}


// ss[0] += sum x^2 (FP64)
__global__ void k_sumsq(const float* x, int64_t n, double* ss)
{
  __shared__ double sh[kThreads];
  double a = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    a += double(x[i]) * double(x[i]);
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(ss, sh[0]);
}

// y = x / sqrt(ss / nss)  (0 when ss = 0)
__global__ void k_normalize(const float* x, const double* ss, int64_t nss, int64_t n, float* y)
{
  const double inv = ss[0] > 0 ? 1.0 / sqrt(ss[0] / double(nss)) : 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    y[i] = float(double(x[i]) * inv);
}

// y = a - b, formed in the inputs' precision and stored as FP32
template <typename TI>
__global__ void k_sub_f(const TI* a, const TI* b, float* y, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    y[i] = float(a[i] - b[i]);
}

// output nodes below which the channel-parallel kernel runs (one thread per
// node would fill fewer than 128 workgroups)
constexpr int64_t kChannelParallelBelow = 32768;

template <int COUT>
hipError_t launch_conv_cp(const ConvArgs& A, bool in_f32, bool out_f32, hipStream_t s)
{
  constexpr int VPB   = kThreads / COUT;
  const int64_t nout  = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const int     cin   = A.ca + A.cb;
  const size_t  shmem = size_t(VPB) * cin * sizeof(float) + size_t(cin) * COUT * sizeof(uint16_t);
  const dim3    grid{unsigned((nout + VPB - 1) / VPB)}, block{unsigned(kThreads)};
  if (in_f32 && out_f32)
    hipLaunchKernelGGL((k_conv_cp<COUT, float, float>), grid, block, shmem, s, A);
  else if (in_f32)
    hipLaunchKernelGGL((k_conv_cp<COUT, float, uint16_t>), grid, block, shmem, s, A);
  else if (out_f32)
    hipLaunchKernelGGL((k_conv_cp<COUT, uint16_t, float>), grid, block, shmem, s, A);
  else
    hipLaunchKernelGGL((k_conv_cp<COUT, uint16_t, uint16_t>), grid, block, shmem, s, A);
  return hipGetLastError();
}

template <int COUT>
hipError_t launch_conv_t(const ConvArgs& A, bool in_f32, bool out_f32, hipStream_t s)
{
  const int64_t nout  = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  if constexpr (COUT >= 16)
    if (nout < kChannelParallelBelow) return launch_conv_cp<COUT>(A, in_f32, out_f32, s);
  const size_t  shmem = size_t(A.ca + A.cb) * COUT * sizeof(uint16_t);
  const dim3    grid{blocks_for(nout)}, block{unsigned(kThreads)};
  if (in_f32 && out_f32)
    hipLaunchKernelGGL((k_conv<COUT, float, float>), grid, block, shmem, s, A);
  else if (in_f32)
    hipLaunchKernelGGL((k_conv<COUT, float, uint16_t>), grid, block, shmem, s, A);
  else if (out_f32)
    hipLaunchKernelGGL((k_conv<COUT, uint16_t, float>), grid, block, shmem, s, A);
  else
    hipLaunchKernelGGL((k_conv<COUT, uint16_t, uint16_t>), grid, block, shmem, s, A);
  return hipGetLastError();
}

template <int COUT>
hipError_t launch_wgrad_t(const WGradArgs& A, bool in_f32, hipStream_t s)
{
  const int64_t nout = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const int64_t runs = (nout + 63) / 64;
  // about 3072 workgroups over the 27 taps, at most 32 runs of 64 nodes each
  const int  chunks = int(std::max<int64_t>(1, std::min<int64_t>(32, runs * 27 / 3072)));
  const dim3 grid{unsigned((runs + chunks - 1) / chunks), 27u}, block{unsigned(kThreads)};
  if (in_f32)
    hipLaunchKernelGGL((k_conv_wgrad<COUT, float>), grid, block, 0, s, A, chunks);
  else
    hipLaunchKernelGGL((k_conv_wgrad<COUT, uint16_t>), grid, block, 0, s, A, chunks);
  return hipGetLastError();
}

}  // namespace

hipError_t conv(const ConvArgs& A, int cout, bool in_f32, bool out_f32, hipStream_t s)
{
  if (A.ca + A.cb > 128) return hipErrorInvalidValue;
  switch (cout) {
    case 3: return launch_conv_t<3>(A, in_f32, out_f32, s);
    case 16: return launch_conv_t<16>(A, in_f32, out_f32, s);
    case 32: return launch_conv_t<32>(A, in_f32, out_f32, s);
    case 64: return launch_conv_t<64>(A, in_f32, out_f32, s);
  }
  return hipErrorInvalidValue;
}

int mfma_kpad(int cin) { return (27 * cin + 31) / 32 * 32; }

bool conv_mfma_supported(const ConvArgs& A, int cout, bool in_f32, bool out_f32)
{
  const int cin = A.ca + A.cb;
  return !in_f32 && !out_f32 && (A.mode == 0 || A.mode == 1) && cin % 8 == 0 && A.ca % 8 == 0 &&
         (cout == 16 || cout == 32 || cout == 64);
}

hipError_t weights_kmajor(const uint16_t* w, int cin, int cout, uint16_t* wt, hipStream_t s)
{
  const int     kpad = mfma_kpad(cin);
  const int64_t n    = int64_t(cout) * kpad;
  hipLaunchKernelGGL(k_weights_kmajor, dim3(std::min<unsigned>(blocks_for(n), 2048)), dim3(kThreads), 0, s, w, cin,
                     cout, kpad, wt);
  return hipGetLastError();
}

hipError_t conv_mfma(const ConvArgs& A, int cout, const uint16_t* wt, hipStream_t s)
{
  if (!conv_mfma_supported(A, cout, false, false)) return hipErrorInvalidValue;
  const int     kpad = mfma_kpad(A.ca + A.cb);
  const int64_t nout = int64_t(A.dout[0]) * A.dout[1] * A.dout[2];
  const dim3 block{unsigned(kThreads)};
  if (nout < kChannelParallelBelow) {
    // few nodes: one 16 x 16 tile per wave
    const int64_t waves = (nout + 15) / 16 * (cout / 16);
    const dim3    g{unsigned((waves + 3) / 4)};
    switch (cout) {
      case 16: hipLaunchKernelGGL(k_conv_mfma_tile<16>, g, block, 0, s, A, wt, kpad); break;
      case 32: hipLaunchKernelGGL(k_conv_mfma_tile<32>, g, block, 0, s, A, wt, kpad); break;
      default: hipLaunchKernelGGL(k_conv_mfma_tile<64>, g, block, 0, s, A, wt, kpad); break;
    }
    return hipGetLastError();
  }
  // nodes per workgroup: 4 waves x RT row tiles of 16
  auto grid = [&](int rt) { return dim3{unsigned((nout + 64 * rt - 1) / (64 * rt))}; };
  switch (cout) {
    case 16: hipLaunchKernelGGL((k_conv_mfma<16, MFMA_RT16>), grid(MFMA_RT16), block, 0, s, A, wt, kpad); break;
    case 32: hipLaunchKernelGGL((k_conv_mfma<32, MFMA_RT32>), grid(MFMA_RT32), block, 0, s, A, wt, kpad); break;
    default: hipLaunchKernelGGL((k_conv_mfma<64, MFMA_RT64>), grid(MFMA_RT64), block, 0, s, A, wt, kpad); break;
  }
  return hipGetLastError();
}

hipError_t conv_wgrad(const WGradArgs& A, int cout, bool in_f32, hipStream_t s)
{
  if (A.ca + A.cb > 128) return hipErrorInvalidValue;
  switch (cout) {
    case 3: return launch_wgrad_t<3>(A, in_f32, s);
    case 16: return launch_wgrad_t<16>(A, in_f32, s);
    case 32: return launch_wgrad_t<32>(A, in_f32, s);
    case 64: return launch_wgrad_t<64>(A, in_f32, s);
  }
  return hipErrorInvalidValue;
}

hipError_t flip_weights(const uint16_t* w, int cin, int cout, int c0, int cpart, uint16_t* wt, hipStream_t s)
{
  hipLaunchKernelGGL(k_flip_weights, dim3(256), dim3(kThreads), 0, s, w, cin, cout, c0, cpart, wt);
  return hipGetLastError();
}
hipError_t relu_bwd(float* g, const uint16_t* y, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_relu_bwd, dim3(std::min<unsigned>(blocks_for(n), 4096)), dim3(kThreads), 0, s, g, y, n);
  return hipGetLastError();
}
hipError_t prolong_bf16(const uint16_t* xc, Dims3 dc, uint16_t* yf, Dims3 df, int C, hipStream_t s)
{
  const int64_t n = int64_t(df.n[0]) * df.n[1] * df.n[2] * C;
  hipLaunchKernelGGL((k_prolong<uint16_t, uint16_t>), dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads),
                     0, s, xc, dc, yf, df, C, 0);
  return hipGetLastError();
}
hipError_t restrict_f32(const float* gf, Dims3 df, float* gc, Dims3 dc, int C, hipStream_t s)
{
  const int64_t n = int64_t(dc.n[0]) * dc.n[1] * dc.n[2] * C;
  hipLaunchKernelGGL(k_restrict, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, gf, df, gc, dc,
                     C);
  return hipGetLastError();
}
hipError_t fill_f32(float* x, float v, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_fill_f, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, x, v, n);
  return hipGetLastError();
}
hipError_t add_f32(float* y, const float* x, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_add_f, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, y, x, n);
  return hipGetLastError();
}
hipError_t loss(const float* y, const float* t, int64_t n, double* sums, float* dy, hipStream_t s)
{
  hipLaunchKernelGGL(k_loss_sums, dim3(std::min<unsigned>(blocks_for(n), 1024)), dim3(kThreads), 0, s, y, t, n, sums);
  hipLaunchKernelGGL(k_loss_grad, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, y, t, n, sums,
                     dy);
  return hipGetLastError();
}
hipError_t adam(float* w, float* m, float* v, float* g, uint16_t* wbf, int64_t n, float lr, float b1, float b2,
                float eps, int step, hipStream_t s)
{
  const float c1 = 1.0f - std::pow(b1, float(step)), c2 = 1.0f - std::pow(b2, float(step));
  hipLaunchKernelGGL(k_adam, dim3(std::min<unsigned>(blocks_for(n), 4096)), dim3(kThreads), 0, s, w, m, v, g, wbf, n,
                     lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}
hipError_t fe_restrict(const FEGrid& G, const double* r, float* rc, hipStream_t s, bool average)
{
  const int64_t n = int64_t(G.nv[0]) * G.nv[1] * G.nv[2] * 3;
  hipLaunchKernelGGL(k_fe_restrict<float>, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, G, r,
                     rc, int(average));
  return hipGetLastError();
}
hipError_t fe_average(const FEGrid& G, const double* x, double* xc, hipStream_t s)
{
  const int64_t n = int64_t(G.nv[0]) * G.nv[1] * G.nv[2] * 3;
  hipLaunchKernelGGL(k_fe_restrict<double>, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, G, x,
                     xc, 1);
  return hipGetLastError();
}
hipError_t fe_prolong_add(const FEGrid& G, const float* ec, const double* ss, int64_t nss, double* z, hipStream_t s,
                          double gain)
{
  const int64_t n = int64_t(G.nn[0]) * G.nn[1] * G.nn[2] * 3;
  hipLaunchKernelGGL(k_fe_prolong_add, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, G, ec, ss,
                     nss, gain, z);
  return hipGetLastError();
}
hipError_t sumsq(const float* x, int64_t n, double* ss, hipStream_t s)
{
  hipLaunchKernelGGL(k_sumsq, dim3(std::min<unsigned>(blocks_for(n), 1024)), dim3(kThreads), 0, s, x, n, ss);
  return hipGetLastError();
}
hipError_t normalize(const float* x, const double* ss, int64_t nss, int64_t n, float* y, hipStream_t s)
{
  hipLaunchKernelGGL(k_normalize, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, x, ss, nss, n,
                     y);
  return hipGetLastError();
}
hipError_t sub_f64(const double* a, const double* b, float* y, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_sub_f<double>, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, a, b, y, n);
  return hipGetLastError();
}
hipError_t to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_to_bf16, dim3(std::min<unsigned>(blocks_for(n), 8192)), dim3(kThreads), 0, s, x, y, n);
  return hipGetLastError();
}

}  // namespace sxu
