// Learned proposal network (host side of unet.hip): a 3-level 3D U-Net on
// BF16 MFMA convolutions with FP32 accumulation and FP32 master weights.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>

namespace tomhip {

// Channels-last implicit-GEMM 3D convolution (zero padding, stride 1).
// X: bf16 [D][H][W][xs], channels [xo, xo + Cin); Wt: bf16 [taps][Cout][Cin]
// (taps 27 = 3x3x3 with tap t = 9(dz+1) + 3(dy+1) + (dx+1), or 1);
// output Yb (bf16, stride ys, offset yo) and/or Yf (f32 [V][Cout]).
// Wm: the same weights in MFMA fragment order (conv_fragments), read
// instead of Wt when taps == 27.
// xq != 0 (3x3x3 only): X is chunk-planar, records of one chunk (32
// channels when Cin is a multiple of 32, else 16; xs = that, xo = 0) with
// chunk q at q xq.
void conv_forward(const uint16_t* X, int xs, int xo, long xq, int Cin, int D, int H, int W, const uint16_t* Wt,
                  const uint16_t* Wm, const float* bias, int Cout, int taps, int relu, uint16_t* Yb, int ys, int yo,
                  float* Yf, hipStream_t s);
// Elements of the fragment-ordered copy of 3x3x3 weights Wt [27][Cout][Cin].
long conv_fragment_elems(int Cout, int Cin);
void conv_fragments(const uint16_t* Wt, int Cout, int Cin, uint16_t* Wm, hipStream_t s);
// dW[t][co][ci] += sum_v dY[v][co] X[v + d_t][ci]; dY bf16 [V][Cout].
// (X chunk-planar when xq != 0, records of xs channels.)
void conv_weight_grad(const uint16_t* X, int xs, int xo, long xq, int Cin, const uint16_t* dY, int Cout, int D, int H,
                      int W, int taps, float* dW, hipStream_t s);

class Unet {
 public:
  // D, H, W: input edges (multiples of 4); nbin orientation bins (<= 76); the
  // head's channels are 1 particle logit, nbin bin logits, 3 centre offsets.
  Unet(hipStream_t s, int D, int H, int W, int nbin, uint64_t seed, bool train);
  ~Unet();
  long num_params() const;
  // all parameters in layer order, each layer W [taps][Cout][Cin] then bias
  void params(float* host, bool get);
  void grads(float* host);
  // Adam state for checkpoints: first and second moments in the parameter
  // order, and the number of updates taken (training networks only)
  void optimizer(float* m, float* v, int64_t* step, bool get);
  // vol: device f32 [D][H][W] (x fastest), already normalised; aux (optional,
  // same layout, normalised): a second input channel
  // head = false stops before the 1x1x1 head (no FP32 logits written)
  void forward(const float* vol, const float* aux, bool head = true);
  double forward_flops() const;
  // tprob: device f32 [V] particle targets in [0, 1]; tbin: device int32 [V]
  // bin index or -1; npos = #{tbin >= 0}; toff (optional): device f32
  // [V][4] offset targets (dx, dy, dz in units of 8 voxels, weight > 0 where
  // supervised), noff = #{weight > 0}. loss_out (host, optional) gets
  // {mean BCE, mean CE, mean offset squared error}.
  void backward(const float* tprob, const int32_t* tbin, const float* toff, float wpos, long npos, long noff,
                double* loss_out);
  void adam(float lr);
  // prob = sigmoid of the particle logit (the logit itself if logit); offs
  // (optional, [V][3]) = the offset channels
  void predict(float* prob, int32_t* bin, bool logit = false, float* offs = nullptr);
  // whole volume vol (device f32 X x Y x Z) and optional aux map of the same
  // size, each normalised here by its global mean and standard deviation, in
  // tiles of the network size with halo voxels of context on each side;
  // prob / bin (device) receive every voxel
  void infer(const float* vol, const float* aux, int X, int Y, int Z, int halo, float* prob, int32_t* bin,
             bool logit = false, float* offs = nullptr);
  const float* logits() const;  // device f32 [V][80]; null before the first forward() with the head
  int D() const;
  int H() const;
  int W() const;

 private:
  void refresh_weights();
  // forward() of the network-size tile at (ox, oy, oz) of a volume X x Y x Z
  // (zero outside it), normalised by the global statistics in sums (device:
  // sum, sum of squares of vol, then of aux; n voxels) when given
  void forward_window(const float* vol, const float* aux, int X, int Y, int Z, int ox, int oy, int oz,
                      const double* sums, long n, bool head);
  // predict()'s outputs from the last decoder activations (the head fused
  // with the prediction), written for the tile core [halo, edge - halo)
  // straight into the X x Y x Z output volumes
  void head_window(int halo, int ox, int oy, int oz, int X, int Y, int Z, float* prob, int32_t* bin, bool logit,
                   float* offs);
  struct Impl;
  std::unique_ptr<Impl> d_;
};

}  // namespace tomhip
