// Learned safety: sample generation from the exact kernels, training with
// BF16 matrix-core kernels (forward, backward, Adam), margin validation and
// the stand-alone inference kernel.  The network is defined in mlp.h; the
// fused forward used by inference and by the transport is in mlp_serve.hpp.
//
// Training works on row-major batches in HBM: features F [B][64], per layer
// pre-activations Z [B][128] (FP32) and activations H [B][128] (BF16).  Every
// product (forward Z = H W^T, backward dH = dZ W and dW = dZ^T H) runs on one
// strided BF16 matrix-core kernel (32x32 tiles per wave, split over K with
// FP32 atomics for the weight gradients, whose K is the batch).
#include <algorithm>
#include <cmath>
#include <vector>
#include "context.h"
#include "mlp.h"
#include "mlp_serve.hpp"
#include "mlp_train.h"
#include "shapes/all_shapes.hpp"
#include "stepping.hpp"

namespace vgm {

// Internal FP32 parameter layout: the public order with W1 padded to 64
// inputs (columns 51..63 stay zero).
constexpr int kIW1 = 0;
constexpr int kIB1 = kIW1 + kMlpHidden * kMlpInPad;
constexpr int kIW2 = kIB1 + kMlpHidden;
constexpr int kIB2 = kIW2 + kMlpHidden * kMlpHidden;
constexpr int kIW3 = kIB2 + kMlpHidden;
constexpr int kIB3 = kIW3 + kMlpHidden * kMlpHidden;
constexpr int kIW4 = kIB3 + kMlpHidden;
constexpr int kIB4 = kIW4 + kMlpHidden * kMlpHidden;
constexpr int kIW5 = kIB4 + kMlpHidden;
constexpr int kIB5 = kIW5 + kMlpHidden;
static_assert(kIB5 + 1 == kMlpInternal, "internal layout");
constexpr int kIW[4] = {kIW1, kIW2, kIW3, kIW4};
constexpr int kIB[4] = {kIB1, kIB2, kIB3, kIB4};

namespace {

constexpr int kH = kMlpHidden;

// ---- sample generation ------------------------------------------------------

struct SampleArgs {
  double c[3], hw[3], inv_hw[3];
  double inv_scale;
  uint64_t seed, first;
  size_t n;
  float *u0, *u1, *u2, *target;  // normalised coordinates, safety / scale
  double *px, *py, *pz, *exact;  // optional: positions and exact safety
};

// Points uniform in the domain box; target = SafetyToOut if inside, else
// SafetyToIn (the safety the transport asks for on the point's side).
template <class S>
__global__ __launch_bounds__(256) void mlp_samples_kernel(const typename S::Data *s, SampleArgs a)
{
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.n) return;
  const uint64_t id = a.first + i;
  double p[3];
  for (int k = 0; k < 3; ++k) p[k] = a.c[k] + a.hw[k] * (2.0 * step_uniform(a.seed, id, uint32_t(k)) - 1.0);
  const V3<double> q{p[0], p[1], p[2]};
  const bool in = S::template contains<double>(*s, q);
  const double saf = in ? S::template safety_to_out<double>(*s, q) : S::template safety_to_in<double>(*s, q);
  a.u0[i] = mlp_coord(p[0], a.c[0], a.inv_hw[0]);
  a.u1[i] = mlp_coord(p[1], a.c[1], a.inv_hw[1]);
  a.u2[i] = mlp_coord(p[2], a.c[2], a.inv_hw[2]);
  a.target[i] = float(saf * a.inv_scale);
  if (a.exact) {
    a.px[i] = p[0];
    a.py[i] = p[1];
    a.pz[i] = p[2];
    a.exact[i] = saf;
  }
}

struct SampleLaunch {
  hipStream_t stream;
  const void *params;
  const SampleArgs &a;
  template <class S>
  void operator()()
  {
    const unsigned blocks = unsigned((a.n + 255) / 256);
    hipLaunchKernelGGL(mlp_samples_kernel<S>, dim3(blocks), dim3(256), 0, stream,
                       static_cast<const typename S::Data *>(params), a);
  }
};

// ---- training kernels -------------------------------------------------------

__global__ __launch_bounds__(256) void mlp_features_kernel(const float *u0, const float *u1, const float *u2,
                                                           int B, __bf16 *F)
{
  const int idx = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (idx >= B * kMlpInPad) return;
  const int i = idx / kMlpInPad, f = idx % kMlpInPad;
  F[idx] = static_cast<__bf16>(mlp_feature(f, u0[i], u1[i], u2[i]));
}

// C[M][N] (row stride ldc) = sum_k A(m,k) B(k,n), A(m,k) = A[m sam + k sak],
// B(k,n) = B[k sbk + n sbn], one of each pair of strides 1.  M, N multiples
// of 64, K and kchunk of 32.  Workgroup b computes the 64 x 64 tile b of C
// over the K chunk blockIdx.y: per 32-deep step the A and B tiles are read
// from global memory along their contiguous dimension (16-byte loads) into
// LDS, k-contiguous, and each of the four waves takes its 32 x 32 quarter in
// two 32x32x16 BF16 MFMAs from LDS.  With `atomic` the tile is added to C
// (split K), otherwise stored.
struct GemmArgs {
  const __bf16 *A, *B;
  float *C;
  int M, N, K, kchunk;
  int sam, sak, sbk, sbn, ldc;
  int atomic;
};

constexpr int kGemmTile = 64, kGemmK = 32, kGemmRow = kGemmK + 8;  // LDS rows padded by 16 bytes

__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs g)
{
  __shared__ __attribute__((aligned(16))) __bf16 As[kGemmTile][kGemmRow];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[kGemmTile][kGemmRow];
  const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
  const int tiles_n = g.N / kGemmTile;
  const int m0 = (int(blockIdx.x) / tiles_n) * kGemmTile, n0 = (int(blockIdx.x) % tiles_n) * kGemmTile;
  const int kb = int(blockIdx.y) * g.kchunk, ke = min(g.K, kb + g.kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // the wave's quarter of the tile
  const int r = lane & 31, h = lane >> 5;
  mlp_f32x16 acc = {};
  for (int k0 = kb; k0 < ke; k0 += kGemmK) {
    if (g.sak == 1) {  // rows of A along k: thread t reads row t / 4, k (t % 4) * 8
      const int m = tid >> 2, kc = (tid & 3) * 8;
      *reinterpret_cast<mlp_bf16x8 *>(&As[m][kc]) =
          *reinterpret_cast<const mlp_bf16x8 *>(g.A + size_t(m0 + m) * size_t(g.sam) + size_t(k0 + kc));
    } else {  // A contiguous along m: thread t reads k = t / 8, m (t % 8) * 8
      const int k = tid >> 3, mc = (tid & 7) * 8;
      const mlp_bf16x8 v =
          *reinterpret_cast<const mlp_bf16x8 *>(g.A + size_t(k0 + k) * size_t(g.sak) + size_t(m0 + mc));
#pragma unroll
      for (int j = 0; j < 8; ++j) As[mc + j][k] = v[j];
    }
    if (g.sbk == 1) {  // columns of B along k
      const int n = tid >> 2, kc = (tid & 3) * 8;
      *reinterpret_cast<mlp_bf16x8 *>(&Bs[n][kc]) =
          *reinterpret_cast<const mlp_bf16x8 *>(g.B + size_t(n0 + n) * size_t(g.sbn) + size_t(k0 + kc));
    } else {  // B contiguous along n
      const int k = tid >> 3, nc = (tid & 7) * 8;
      const mlp_bf16x8 v =
          *reinterpret_cast<const mlp_bf16x8 *>(g.B + size_t(k0 + k) * size_t(g.sbk) + size_t(n0 + nc));
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[nc + j][k] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGemmK; kk += 16) {
      const mlp_bf16x8 a = *reinterpret_cast<const mlp_bf16x8 *>(&As[wm + r][kk + 8 * h]);
      const mlp_bf16x8 b = *reinterpret_cast<const mlp_bf16x8 *>(&Bs[wn + r][kk + 8 * h]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    float *c = g.C + size_t(m0 + wm + 8 * (v / 4) + 4 * h + (v % 4)) * size_t(g.ldc) + size_t(n0 + wn + r);
    if (g.atomic)
      atomicAdd(c, acc[v]);
    else
      *c = acc[v];
  }
}

// Z += bias (in place, the pre-activation kept for the backward pass);
// H = BF16(SiLU(Z)).
__global__ __launch_bounds__(256) void mlp_bias_act_kernel(float *Z, const float *bias, __bf16 *H, int n)
{
  const int idx = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (idx >= n) return;
  const float z = Z[idx] + bias[idx % kH];
  Z[idx] = z;
  H[idx] = static_cast<__bf16>(mlp_silu(z));
}

// Output layer and loss.  y = w5 . H4 + b5; e = y - t; loss = mean(w e^2)
// with w = over_weight for overestimates (e > 0), 1 otherwise; dy = dloss/dy.
__global__ __launch_bounds__(256) void mlp_loss_kernel(const __bf16 *H4, const float *w5, const float *b5,
                                                       const float *target, int B, float over_weight, float *y,
                                                       float *dy, float *loss)
{
  const int i = int(blockIdx.x) * 256 + int(threadIdx.x);
  float l = 0.0f;
  if (i < B) {
    float s = 0.0f;
    for (int r = 0; r < kH; ++r) s += w5[r] * static_cast<float>(H4[size_t(i) * kH + r]);
    const float yy = s + b5[0];
    const float e = yy - target[i];
    const float w = e > 0.0f ? over_weight : 1.0f;
    if (y) y[i] = yy;
    dy[i] = 2.0f * w * e / float(B);
    l = w * e * e / float(B);
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63u) == 0 && loss) atomicAdd(loss, l);
}

// Backward through one SiLU layer: g = dH[i][r] (or dy[i] w5[r] for the
// last hidden layer), dZ = g SiLU'(Z), stored in BF16 for the products.  A
// workgroup takes kBwdRows rows, two halves of 128 threads (one per unit)
// over half of them each; the per-unit sums of dZ (the bias gradient) and,
// for the last hidden layer, of dy H4 and dy (dw5, db5) are combined in LDS
// and written as the workgroup's partial, which mlp_bwd_reduce_kernel adds
// up: atomics of every workgroup on the same 128 addresses serialise.
constexpr int kBwdRows = 32;
__global__ __launch_bounds__(256) void mlp_bwd_act_kernel(const float *dH, const float *dy, const float *w5,
                                                          const __bf16 *H4, const float *Z, __bf16 *dZ,
                                                          float *part, int B)
{
  __shared__ float red[3][kH];
  const int r = int(threadIdx.x) & (kH - 1), half = int(threadIdx.x) >> 7;
  const int i0 = int(blockIdx.x) * kBwdRows + half * (kBwdRows / 2);
  float sb = 0.0f, sw = 0.0f, sb5 = 0.0f;
#pragma unroll
  for (int j = 0; j < kBwdRows / 2; ++j) {
    const int i = i0 + j;
    if (i >= B) break;
    const size_t idx = size_t(i) * kH + r;
    const float g = dH ? dH[idx] : dy[i] * w5[r];
    const float z = Z[idx];
    const float sg = 1.0f / (1.0f + __expf(-z));
    const float d = g * sg * (1.0f + z * (1.0f - sg));
    dZ[idx] = static_cast<__bf16>(d);
    sb += d;
    if (!dH) {
      sw += dy[i] * static_cast<float>(H4[idx]);
      sb5 += dy[i];
    }
  }
  if (half) {
    red[0][r] = sb;
    red[1][r] = sw;
    red[2][r] = sb5;
  }
  __syncthreads();
  if (!half) {
    float *pb = part + size_t(blockIdx.x) * (2 * kH + 1);
    pb[r] = sb + red[0][r];
    if (!dH) {
      pb[kH + r] = sw + red[1][r];
      if (r == 0) pb[2 * kH] = sb5 + red[2][0];
    }
  }
}

// The bias gradient (and for the last hidden layer dw5, db5) as the sum of
// the nblk workgroup partials of mlp_bwd_act_kernel: eight groups of 128
// threads (one per unit) each add every eighth partial, with independent
// loads in flight, and the groups are combined in LDS.
constexpr int kBwdReduceGroups = 8;
__global__ __launch_bounds__(kBwdReduceGroups * kH) void mlp_bwd_reduce_kernel(const float *part, int nblk,
                                                                               float *db, float *dw5, float *db5)
{
  __shared__ float red[kBwdReduceGroups][3][kH];
  const int r = int(threadIdx.x) & (kH - 1), grp = int(threadIdx.x) / kH;
  float sb = 0.0f, sw = 0.0f, sb5 = 0.0f;
#pragma unroll 4
  for (int b = grp; b < nblk; b += kBwdReduceGroups) {
    const float *pb = part + size_t(b) * (2 * kH + 1);
    sb += pb[r];
    if (dw5) {
      sw += pb[kH + r];
      sb5 += pb[2 * kH];
    }
  }
  red[grp][0][r] = sb;
  red[grp][1][r] = sw;
  red[grp][2][r] = sb5;
  __syncthreads();
  if (grp == 0) {
    for (int k = 1; k < kBwdReduceGroups; ++k) {
      sb += red[k][0][r];
      sw += red[k][1][r];
      sb5 += red[k][2][r];
    }
    db[r] = sb;
    if (dw5) {
      dw5[r] = sw;
      if (r == 0) *db5 = sb5;
    }
  }
}

// Adam step on the internal parameters; refreshes their BF16 copies.
__global__ __launch_bounds__(256) void mlp_adam_kernel(float *p, const float *g, float *m, float *v, __bf16 *pb,
                                                       int n, float lr, float c1, float c2)
{
  const int i = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
  p[i] = pi;
  pb[i] = static_cast<__bf16>(pi);
}

__global__ __launch_bounds__(256) void mlp_to_bf16_kernel(const float *p, __bf16 *pb, int n)
{
  const int i = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (i < n) pb[i] = static_cast<__bf16>(p[i]);
}

// Serving image from the internal parameters (padding, BF16).
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float *p, MlpImage *img)
{
  const int idx = int(blockIdx.x) * 256 + int(threadIdx.x);
  const int n1 = kH * kMlpRow1, nh = kH * kMlpRowH;
  if (idx < n1) {
    const int m = idx / kMlpRow1, k = idx % kMlpRow1;
    img->w1[m][k] = static_cast<__bf16>(k < kMlpInPad ? p[kIW1 + m * kMlpInPad + k] : 0.0f);
    return;
  }
  int j = idx - n1;
  if (j < 3 * nh) {
    const int l = j / nh;
    j %= nh;
    const int m = j / kMlpRowH, k = j % kMlpRowH;
    float w = 0.0f;
    if (k < kH) w = p[kIW[l + 1] + m * kH + k];
    __bf16 *row = (l == 0 ? img->w2[m] : (l == 1 ? img->w3[m] : img->w4[m]));
    row[k] = static_cast<__bf16>(w);
    return;
  }
  j -= 3 * nh;
  if (j < 5 * kH) {
    const int l = j / kH, r = j % kH;
    if (l < 4)
      (l == 0 ? img->b1 : (l == 1 ? img->b2 : (l == 2 ? img->b3 : img->b4)))[r] = p[kIB[l] + r];
    else
      img->w5[r] = p[kIW5 + r];
    return;
  }
  j -= 5 * kH;
  if (j == 0) {
    img->b5 = p[kIB5];
    img->pad[0] = img->pad[1] = img->pad[2] = 0.0f;
  }
}

// Stand-alone inference: points in the shape's frame, one wave per 64
// points.  bound: max(0, y scale - margin).
__global__ __launch_bounds__(256) void mlp_infer_kernel(MlpDesc d, const double *x, const double *y,
                                                        const double *z, size_t n, double *out, int bound)
{
  __shared__ mlp_act_rows act[4];
  const int lane = int(threadIdx.x & 63u);
  const size_t wave = size_t(blockIdx.x) * 4 + (threadIdx.x >> 6), nw = size_t(gridDim.x) * 4;
  for (size_t base = wave * 64; base < n; base += nw * 64) {
    const size_t i = base + size_t(lane);
    const bool act_lane = i < n;
    const float ux = act_lane ? mlp_coord(x[i], d.c[0], d.inv_hw[0]) : 0.0f;
    const float uy = act_lane ? mlp_coord(y[i], d.c[1], d.inv_hw[1]) : 0.0f;
    const float uz = act_lane ? mlp_coord(z[i], d.c[2], d.inv_hw[2]) : 0.0f;
    const float v = mlp_eval_wave(*d.img, act[threadIdx.x >> 6], ux, uy, uz);
    if (act_lane) {
      const double pred = double(v) * d.scale;
      out[i] = bound ? fmax(0.0, pred - d.margin) : pred;
    }
  }
}

unsigned blocks_for(size_t n, int per) { return unsigned((n + size_t(per) - 1) / size_t(per)); }

// Deterministic initialisation (uniform +-1/sqrt(fan_in), as torch's Linear).
std::vector<float> init_params(uint64_t seed)
{
  std::vector<float> p(kMlpInternal, 0.0f);
  uint64_t s = seed ^ 0x243F6A8885A308D3ull;
  auto uni = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    return double(mix64(s) >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int rows, int cols, int ld, int fan_in) {
    const double a = 1.0 / std::sqrt(double(fan_in));
    for (int r = 0; r < rows; ++r)
      for (int k = 0; k < cols; ++k) p[size_t(off + r * ld + k)] = float((2.0 * uni() - 1.0) * a);
  };
  fill(kIW1, kH, kMlpIn, kMlpInPad, kMlpIn);
  fill(kIB1, 1, kH, kH, kMlpIn);
  for (int l = 1; l < 4; ++l) {
    fill(kIW[l], kH, kH, kH, kH);
    fill(kIB[l], 1, kH, kH, kH);
  }
  fill(kIW5, 1, kH, kH, kH);
  fill(kIB5, 1, 1, 1, kH);
  return p;
}

}  // namespace

// ---- host side --------------------------------------------------------------

void mlp_public_to_internal(const float *pub, float *in)
{
  for (int m = 0; m < kH; ++m)
    for (int k = 0; k < kMlpInPad; ++k) in[kIW1 + m * kMlpInPad + k] = k < kMlpIn ? pub[kMlpOffW1 + m * kMlpIn + k] : 0.0f;
  std::copy(pub + kMlpOffB1, pub + kMlpParams, in + kIB1);
}

void mlp_internal_to_public(const float *in, float *pub)
{
  for (int m = 0; m < kH; ++m)
    for (int k = 0; k < kMlpIn; ++k) pub[kMlpOffW1 + m * kMlpIn + k] = in[kIW1 + m * kMlpInPad + k];
  std::copy(in + kIB1, in + kMlpInternal, pub + kMlpOffB1);
}

namespace {

hipError_t gemm(hipStream_t st, const __bf16 *A, int sam, int sak, const __bf16 *B, int sbk, int sbn, float *C,
                int ldc, int M, int N, int K, int kchunk, bool atomic)
{
  GemmArgs g{A, B, C, M, N, K, kchunk, sam, sak, sbk, sbn, ldc, atomic ? 1 : 0};
  const int tiles = (M / kGemmTile) * (N / kGemmTile);
  hipLaunchKernelGGL(mlp_gemm_kernel, dim3(unsigned(tiles), unsigned((K + kchunk - 1) / kchunk)), dim3(256), 0, st,
                     g);
  return hipGetLastError();
}

// Device work buffers of one training run (batch B).
struct Work {
  char *buf = nullptr;
  __bf16 *F = nullptr, *H[4] = {}, *dZ[2] = {}, *pb = nullptr;
  float *Z[4] = {}, *dH = nullptr, *y = nullptr, *dy = nullptr, *loss = nullptr;
  float *grad = nullptr, *m = nullptr, *v = nullptr;
  float *part = nullptr;  // mlp_bwd_act_kernel partials: per workgroup 2 kH + 1
  ~Work()
  {
    if (buf) (void)hipFree(buf);
  }
  hipError_t alloc(int B)
  {
    const size_t bh = align256(size_t(B) * kH * 2), bz = align256(size_t(B) * kH * 4);
    const size_t bf = align256(size_t(B) * kMlpInPad * 2), bb = align256(size_t(B) * 4);
    const size_t bp = align256(size_t(kMlpInternal) * 4), bpb = align256(size_t(kMlpInternal) * 2);
    const size_t bpart = align256(size_t((B + kBwdRows - 1) / kBwdRows) * (2 * kH + 1) * 4);
    const size_t total = bf + 4 * bh + 2 * bh + 4 * bz + bz + 2 * bb + 256 + 3 * bp + bpb + bpart;
    hipError_t e = hipMalloc(&buf, total);
    if (e != hipSuccess) return e;
    char *q = buf;
    auto take = [&](size_t b) {
      char *r = q;
      q += b;
      return r;
    };
    F = reinterpret_cast<__bf16 *>(take(bf));
    for (auto &h : H) h = reinterpret_cast<__bf16 *>(take(bh));
    for (auto &d : dZ) d = reinterpret_cast<__bf16 *>(take(bh));
    for (auto &z : Z) z = reinterpret_cast<float *>(take(bz));
    dH = reinterpret_cast<float *>(take(bz));
    y = reinterpret_cast<float *>(take(bb));
    dy = reinterpret_cast<float *>(take(bb));
    loss = reinterpret_cast<float *>(take(256));
    part = reinterpret_cast<float *>(take(bpart));
    grad = reinterpret_cast<float *>(take(bp));
    m = reinterpret_cast<float *>(take(bp));
    v = reinterpret_cast<float *>(take(bp));
    pb = reinterpret_cast<__bf16 *>(take(bpb));
    return hipSuccess;
  }
};

// Forward and backward of one batch: gradients of the mean loss into w.grad
// (zeroed here), the loss added to *w.loss.
hipError_t forward_backward(hipStream_t st, Work &w, const float *params, const float *u0, const float *u1,
                            const float *u2, const float *target, int B, float over_weight)
{
  hipError_t e;
  hipLaunchKernelGGL(mlp_features_kernel, dim3(blocks_for(size_t(B) * kMlpInPad, 256)), dim3(256), 0, st, u0, u1,
                     u2, B, w.F);
  const __bf16 *in = w.F;
  int ld = kMlpInPad;
  for (int l = 0; l < 4; ++l) {
    // Z = H W^T: A(i, k) = H[i][k]; B(k, n) = W[n][k]
    if ((e = gemm(st, in, ld, 1, w.pb + kIW[l], 1, ld, w.Z[l], kH, B, kH, ld, ld, false)) != hipSuccess) return e;
    hipLaunchKernelGGL(mlp_bias_act_kernel, dim3(blocks_for(size_t(B) * kH, 256)), dim3(256), 0, st, w.Z[l],
                       params + kIB[l], w.H[l], B * kH);
    in = w.H[l];
    ld = kH;
  }
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(blocks_for(size_t(B), 256)), dim3(256), 0, st, w.H[3],
                     params + kIW5, params + kIB5, target, B, over_weight, w.y, w.dy, w.loss);
  if ((e = hipMemsetAsync(w.grad, 0, size_t(kMlpInternal) * sizeof(float), st)) != hipSuccess) return e;
  const int nblk = (B + kBwdRows - 1) / kBwdRows;
  hipLaunchKernelGGL(mlp_bwd_act_kernel, dim3(unsigned(nblk)), dim3(2 * kH), 0, st, nullptr, w.dy, params + kIW5,
                     w.H[3], w.Z[3], w.dZ[1], w.part, B);
  hipLaunchKernelGGL(mlp_bwd_reduce_kernel, dim3(1), dim3(kBwdReduceGroups * kH), 0, st, w.part, nblk, w.grad + kIB[3],
                     w.grad + kIW5,
                     w.grad + kIB5);
  const int kchunk = std::min(B, 512);
  for (int l = 3; l >= 0; --l) {
    const __bf16 *dz = w.dZ[l & 1];
    const __bf16 *hin = l ? w.H[l - 1] : w.F;
    const int nin = l ? kH : kMlpInPad;
    // dW = dZ^T H: A(m = unit, k = i) = dZ[i][m]; B(k = i, n) = H[i][n]
    if ((e = gemm(st, dz, 1, kH, hin, nin, 1, w.grad + kIW[l], nin, kH, nin, B, kchunk, true)) != hipSuccess)
      return e;
    if (l == 0) break;
    // dH = dZ W: A(i, k = unit) = dZ[i][k]; B(k = unit, n) = W[k][n]
    if ((e = gemm(st, dz, kH, 1, w.pb + kIW[l], kH, 1, w.dH, kH, B, kH, kH, kH, false)) != hipSuccess) return e;
    hipLaunchKernelGGL(mlp_bwd_act_kernel, dim3(unsigned(nblk)), dim3(2 * kH), 0, st, w.dH, nullptr, nullptr, nullptr,
                       w.Z[l - 1], w.dZ[(l - 1) & 1], w.part, B);
    hipLaunchKernelGGL(mlp_bwd_reduce_kernel, dim3(1), dim3(kBwdReduceGroups * kH), 0, st, w.part, nblk,
                       w.grad + kIB[l - 1], nullptr, nullptr);
  }
  return hipGetLastError();
}

int state_for(vgm_ctx *c, int shape, MlpState *&st)
{
  if (shape < 0 || size_t(shape) >= c->host.shapes.size()) return fail(VGM_EINVAL, "invalid shape id");
  if (c->mlps.size() < c->host.shapes.size()) c->mlps.resize(c->host.shapes.size());
  st = &c->mlps[size_t(shape)];
  if (!st->params) VGM_HIP(hipMalloc(&st->params, size_t(kMlpInternal) * sizeof(float)));
  if (!st->img) VGM_HIP(hipMalloc(&st->img, sizeof(MlpImage)));
  return VGM_OK;
}

void set_domain(MlpState &st, const ShapeRec &rec, double domain)
{
  double d2 = 0;
  for (int k = 0; k < 3; ++k) {
    st.c[k] = 0.5 * (rec.bmin[k] + rec.bmax[k]);
    st.hw[k] = domain * 0.5 * (rec.bmax[k] - rec.bmin[k]);
    if (!(st.hw[k] > 0)) st.hw[k] = 1.0;
    d2 += (rec.bmax[k] - rec.bmin[k]) * (rec.bmax[k] - rec.bmin[k]);
  }
  st.scale = std::sqrt(d2) > 0 ? std::sqrt(d2) : 1.0;
}

SampleArgs sample_args(const MlpState &st, uint64_t seed, uint64_t first, size_t n)
{
  SampleArgs a{};
  for (int k = 0; k < 3; ++k) {
    a.c[k] = st.c[k];
    a.hw[k] = st.hw[k];
    a.inv_hw[k] = 1.0 / st.hw[k];
  }
  a.inv_scale = 1.0 / st.scale;
  a.seed = seed;
  a.first = first;
  a.n = n;
  return a;
}

int launch_samples(vgm_ctx *c, int shape, const SampleArgs &a)
{
  const ShapeRec &rec = c->host.shapes[size_t(shape)];
  SampleLaunch L{c->stream, c->dev.pool + rec.off, a};
  if (!dispatch_shape_type(rec.type, L)) return fail(VGM_EINVAL, "unknown shape type");
  VGM_HIP(hipGetLastError());
  return VGM_OK;
}

}  // namespace

MlpDesc mlp_desc(const MlpState &st)
{
  MlpDesc d{};
  d.img = st.img;
  for (int k = 0; k < 3; ++k) {
    d.c[k] = st.c[k];
    d.inv_hw[k] = 1.0 / st.hw[k];
  }
  d.scale = st.scale;
  d.margin = st.margin;
  return d;
}

int mlp_pack(vgm_ctx *c, MlpState &st)
{
  const size_t n = size_t(kH) * kMlpRow1 + 3 * size_t(kH) * kMlpRowH + 5 * kH + 1;
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, st.params, st.img);
  VGM_HIP(hipGetLastError());
  return VGM_OK;
}

int mlp_infer(vgm_ctx *c, const MlpState &st, const double *x, const double *y, const double *z, size_t n,
              double *out, bool bound)
{
  if (n == 0) return VGM_OK;
  const unsigned grid = std::max(1u, std::min(unsigned(c->num_cus) * 8u, blocks_for(n, 256)));
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(256), 0, c->stream, mlp_desc(st), x, y, z, n, out,
                     bound ? 1 : 0);
  VGM_HIP(hipGetLastError());
  return VGM_OK;
}

int mlp_train(vgm_ctx *c, int shape, const vgm_mlp_config &cfg, vgm_mlp_report &rep)
{
  const int B = cfg.batch;
  if (B < 256 || B % 256 || cfg.samples < B || cfg.epochs < 1 || cfg.validation < 1 || !(cfg.domain >= 1.0))
    return fail(VGM_EINVAL, "mlp config: batch must be a multiple of 256, samples >= batch, epochs >= 1, "
                            "validation >= 1, domain >= 1");
  int rc = sync_geometry(c);
  if (rc) return rc;
  MlpState *st = nullptr;
  if ((rc = state_for(c, shape, st))) return rc;
  const ShapeRec &rec = c->host.shapes[size_t(shape)];
  set_domain(*st, rec, cfg.domain);
  hipStream_t s = c->stream;
  hipEvent_t e0, e1;
  VGM_HIP(hipEventCreate(&e0));
  VGM_HIP(hipEventCreate(&e1));
  const size_t N = size_t(cfg.samples) / size_t(B) * size_t(B);
  Work w;
  float *samples = nullptr;
  double *val = nullptr;
  auto done = [&](int code) {
    (void)hipStreamSynchronize(s);
    if (samples) (void)hipFree(samples);
    if (val) (void)hipFree(val);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return code;
  };
  if (w.alloc(B) != hipSuccess || hipMalloc(&samples, 4 * N * sizeof(float)) != hipSuccess)
    return done(fail(VGM_ENOMEM, "mlp training buffers"));
  SampleArgs sa = sample_args(*st, cfg.seed, 0, N);
  sa.u0 = samples;
  sa.u1 = samples + N;
  sa.u2 = samples + 2 * N;
  sa.target = samples + 3 * N;
  (void)hipEventRecord(e0, s);
  if ((rc = launch_samples(c, shape, sa))) return done(rc);
  const std::vector<float> p0 = init_params(cfg.seed);
  if (hipMemcpyAsync(st->params, p0.data(), p0.size() * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(w.m, 0, size_t(kMlpInternal) * 4, s) != hipSuccess ||
      hipMemsetAsync(w.v, 0, size_t(kMlpInternal) * 4, s) != hipSuccess)
    return done(fail(VGM_EHIP, "mlp init copy"));
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3(blocks_for(kMlpInternal, 256)), dim3(256), 0, s, st->params, w.pb,
                     kMlpInternal);
  const int64_t per_epoch = int64_t(N / size_t(B));
  const int64_t steps = per_epoch * cfg.epochs;
  std::vector<float> losses;
  for (int64_t t = 0; t < steps; ++t) {
    const size_t off = size_t(t % per_epoch) * size_t(B);
    const bool last = t >= steps - std::min<int64_t>(steps, 16);
    if (last && hipMemsetAsync(w.loss, 0, sizeof(float), s) != hipSuccess) return done(fail(VGM_EHIP, "memset"));
    if (forward_backward(s, w, st->params, sa.u0 + off, sa.u1 + off, sa.u2 + off, sa.target + off, B,
                         cfg.over_weight) != hipSuccess)
      return done(fail(VGM_EHIP, "mlp training step launch"));
    const double frac = double(t) / double(steps);
    const float lr = float(cfg.lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac))));
    const float c1 = float(1.0 / (1.0 - std::pow(0.9, double(t + 1))));
    const float c2 = float(1.0 / (1.0 - std::pow(0.999, double(t + 1))));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(blocks_for(kMlpInternal, 256)), dim3(256), 0, s, st->params, w.grad,
                       w.m, w.v, w.pb, kMlpInternal, lr, c1, c2);
    if (last) {
      float l = 0;
      if (hipMemcpyAsync(&l, w.loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return done(fail(VGM_EHIP, "mlp loss readback"));
      losses.push_back(l);
    }
  }
  (void)hipEventRecord(e1, s);
  if ((rc = mlp_pack(c, *st))) return done(rc);
  // validation on fresh points: largest overestimate of the served function
  const size_t V = size_t(cfg.validation);
  if (hipMalloc(&val, 5 * V * sizeof(double) + 4 * V * sizeof(float)) != hipSuccess)
    return done(fail(VGM_ENOMEM, "mlp validation buffers"));
  SampleArgs va = sample_args(*st, cfg.seed ^ 0x5DEECE66Dull, uint64_t(1) << 40, V);
  va.px = val;
  va.py = val + V;
  va.pz = val + 2 * V;
  va.exact = val + 3 * V;
  float *vf = reinterpret_cast<float *>(val + 5 * V);
  va.u0 = vf;
  va.u1 = vf + V;
  va.u2 = vf + 2 * V;
  va.target = vf + 3 * V;
  if ((rc = launch_samples(c, shape, va))) return done(rc);
  st->margin = 0.0;
  if ((rc = mlp_infer(c, *st, va.px, va.py, va.pz, V, val + 4 * V, false))) return done(rc);
  std::vector<double> hv(2 * V);
  if (hipMemcpyAsync(hv.data(), val + 3 * V, 2 * V * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(VGM_EHIP, "mlp validation readback"));
  double max_over = -1e300, sum_abs = 0, sum_sq = 0;
  for (size_t i = 0; i < V; ++i) {
    const double err = hv[V + i] - hv[i];
    max_over = std::max(max_over, err);
    sum_abs += std::fabs(err);
    sum_sq += err * err;
  }
  st->margin = cfg.margin_factor * std::max(0.0, max_over) + cfg.margin_abs * st->scale;
  st->ready = true;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  rep = vgm_mlp_report{};
  rep.margin = st->margin;
  rep.max_over = max_over;
  rep.mean_abs_err = sum_abs / double(V);
  rep.rms_err = std::sqrt(sum_sq / double(V));
  double lsum = 0;
  for (float l : losses) lsum += l;
  rep.final_loss = losses.empty() ? 0.0 : lsum / double(losses.size());
  rep.train_ms = ms;
  rep.steps = steps;
  rep.samples = int64_t(N);
  rep.scale = st->scale;
  rep.gflop = 6.0 * double(kMlpParams) * double(N) * double(cfg.epochs) * 1e-9;
  return done(VGM_OK);
}

int mlp_gradient(vgm_ctx *c, const float *params_pub, const float *u, const float *target, size_t n,
                 float over_weight, float *grad_pub, float *loss)
{
  if (n == 0 || n % 64) return fail(VGM_EINVAL, "gradient batch must be a positive multiple of 64");
  if (n > (size_t(1) << 24)) return fail(VGM_EINVAL, "gradient batch too large");
  hipStream_t s = c->stream;
  const int B = int(n);
  Work w;
  float *d = nullptr;
  auto done = [&](int code) {
    (void)hipStreamSynchronize(s);
    if (d) (void)hipFree(d);
    return code;
  };
  if (w.alloc(B) != hipSuccess || hipMalloc(&d, (size_t(kMlpInternal) + 4 * n) * sizeof(float)) != hipSuccess)
    return done(fail(VGM_ENOMEM, "mlp gradient buffers"));
  std::vector<float> pin(kMlpInternal), gin(kMlpInternal), hu(3 * n);
  mlp_public_to_internal(params_pub, pin.data());
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) hu[size_t(k) * n + i] = u[3 * i + size_t(k)];
  float *dp = d, *du = d + kMlpInternal, *dt = du + 3 * n;
  if (hipMemcpyAsync(dp, pin.data(), pin.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(du, hu.data(), hu.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dt, target, n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(w.loss, 0, sizeof(float), s) != hipSuccess)
    return done(fail(VGM_EHIP, "mlp gradient upload"));
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3(blocks_for(kMlpInternal, 256)), dim3(256), 0, s, dp, w.pb,
                     kMlpInternal);
  if (forward_backward(s, w, dp, du, du + n, du + 2 * n, dt, B, over_weight) != hipSuccess)
    return done(fail(VGM_EHIP, "mlp gradient launch"));
  if (hipMemcpyAsync(gin.data(), w.grad, gin.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(loss, w.loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(VGM_EHIP, "mlp gradient readback"));
  mlp_internal_to_public(gin.data(), grad_pub);
  return done(VGM_OK);
}

int mlp_set(vgm_ctx *c, int shape, const float *params_pub, const double *geom8)
{
  MlpState *st = nullptr;
  int rc = state_for(c, shape, st);
  if (rc) return rc;
  std::vector<float> pin(kMlpInternal);
  mlp_public_to_internal(params_pub, pin.data());
  VGM_HIP(hipMemcpyAsync(st->params, pin.data(), pin.size() * 4, hipMemcpyHostToDevice, c->stream));
  for (int k = 0; k < 3; ++k) {
    st->c[k] = geom8[k];
    st->hw[k] = geom8[3 + k];
  }
  st->scale = geom8[6];
  st->margin = geom8[7];
  if ((rc = mlp_pack(c, *st))) return rc;
  VGM_HIP(hipStreamSynchronize(c->stream));
  st->ready = true;
  return VGM_OK;
}

int mlp_get(vgm_ctx *c, int shape, float *params_pub, double *geom8)
{
  if (shape < 0 || size_t(shape) >= c->mlps.size() || !c->mlps[size_t(shape)].ready)
    return fail(VGM_ESTATE, "no trained network for this shape");
  const MlpState &st = c->mlps[size_t(shape)];
  std::vector<float> pin(kMlpInternal);
  VGM_HIP(hipMemcpyAsync(pin.data(), st.params, pin.size() * 4, hipMemcpyDeviceToHost, c->stream));
  VGM_HIP(hipStreamSynchronize(c->stream));
  mlp_internal_to_public(pin.data(), params_pub);
  for (int k = 0; k < 3; ++k) {
    geom8[k] = st.c[k];
    geom8[3 + k] = st.hw[k];
  }
  geom8[6] = st.scale;
  geom8[7] = st.margin;
  return VGM_OK;
}

void mlp_release(vgm_ctx *c)
{
  for (MlpState &st : c->mlps) {
    if (st.params) (void)hipFree(st.params);
    if (st.img) (void)hipFree(st.img);
    st = MlpState{};
  }
}

}  // namespace vgm
