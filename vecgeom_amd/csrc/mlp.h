// Learned safety: a small multilayer perceptron per solid that predicts the
// safety (distance to the surface from inside or outside) at a point, served
// as a conservative lower bound after subtracting a validated margin.
//
// Network (fixed shape):
//   u      = (p - c) / hw per axis, c / hw the centre and half-widths of the
//            solid's domain box (1.3 x its bounding box), so u is in [-1, 1]
//   f      = [u, sin(2^k pi u), cos(2^k pi u) for k = 0..7]   51 features
//            (feature 3 + 6k + a = sin of axis a, 3 + 6k + 3 + a = cos)
//   h1..h4 = SiLU(W h + b), 128 units each (h0 = f)
//   y      = w5 . h4 + b5, prediction = y * scale (scale = bounding-box diagonal)
// Matrix products take BF16 operands (features and activations rounded to
// BF16, weights kept in FP32 and rounded to BF16 for the products) and
// accumulate in FP32; biases, SiLU and the output layer are FP32.
//
// Public parameter order (FP32, kMlpParams values): W1[128][51], b1[128],
// W2[128][128], b2, W3[128][128], b3, W4[128][128], b4, w5[128], b5.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgm {

constexpr int kMlpFreqs = 8;
constexpr int kMlpIn = 3 + 6 * kMlpFreqs;  // 51
constexpr int kMlpInPad = 64;
constexpr int kMlpHidden = 128;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHidden * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW4 = kMlpOffB3 + kMlpHidden;
constexpr int kMlpOffB4 = kMlpOffW4 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW5 = kMlpOffB4 + kMlpHidden;
constexpr int kMlpOffB5 = kMlpOffW5 + kMlpHidden;
constexpr int kMlpParams = kMlpOffB5 + 1;  // 56321
static_assert(kMlpParams == 56321, "parameter count");

// Rows of the weight images are padded by 8 BF16 (16 bytes) so that the 32
// rows one matrix instruction reads fall in different banks / channels.
constexpr int kMlpRow1 = kMlpInPad + 8;
constexpr int kMlpRowH = kMlpHidden + 8;

// Serving image of one network: BF16 weights with padded rows and FP32
// biases, the layout the fused forward (mlp_serve.hpp) reads.
struct MlpImage {
  __bf16 w1[kMlpHidden][kMlpRow1];
  __bf16 w2[kMlpHidden][kMlpRowH];
  __bf16 w3[kMlpHidden][kMlpRowH];
  __bf16 w4[kMlpHidden][kMlpRowH];
  float b1[kMlpHidden], b2[kMlpHidden], b3[kMlpHidden], b4[kMlpHidden];
  float w5[kMlpHidden];
  float b5, pad[3];
};
static_assert(sizeof(MlpImage) % 16 == 0, "image is whole 16-byte chunks");

// Per-network constants of the served bound.
struct MlpDesc {
  const MlpImage *img;  // device image
  double c[3];          // domain box centre
  double inv_hw[3];     // 1 / domain half-widths
  double scale;         // prediction = y * scale
  double margin;        // served bound = prediction - margin
};

// Host-side state of the network of one shape in a context.  Parameters
// are kept in an internal FP32 layout that pads W1 to 64 inputs
// (kMlpInternal values, see mlp.hip); the device image is rebuilt from them.
constexpr int kMlpInternal = kMlpParams + kMlpHidden * (kMlpInPad - kMlpIn);
struct MlpState {
  float *params = nullptr;   // device, internal layout
  MlpImage *img = nullptr;   // device serving image
  double c[3] = {0, 0, 0}, hw[3] = {1, 1, 1};
  double scale = 1.0, margin = 0.0;
  bool ready = false;
};

}  // namespace vgm
