// mlp.hip: the learned scheduler of the BFS/CC repair.
//
// A 12-64-64-3 perceptron (ReLU) with bf16 weights and fp32 accumulation. Per deleted parent
// edge it predicts, from int-derived features of the child endpoint c:
//   out[0]  logit of "c ends unreached", i.e. the deletion splits vertex 0's component;
//   out[1]  log1p of the number of vertices the deletion's cascade invalidates;
//   out[2]  the number of levels that cascade spans.
// The repair uses the predictions only to size and order work (rounds queued before the first
// read-back, local repair versus full recompute); every answer is still computed by the plain
// search, so results are bit-identical with and without the network.
//
// Training data are logged by the plain repair itself: features of up to kMlpSample candidates
// per batch, and after the repair their outcomes (root attribution: each invalidated vertex
// is charged to the candidate whose cascade reached it). Training runs on the device with the
// kernels below: a wave per slice of 64 samples runs forward and backward with lane j owning
// hidden unit j (its weights in registers, the other units' activations by lane shuffles) and
// sums the slice's weight gradients in registers; the slices are then reduced in a fixed
// order (deterministic) and Adam updates fp32 master weights, which are rounded to
// nearest-even bf16 for the forward pass after every step.
#include <cmath>
#include <vector>

#include "../host/rng.h"
#include "common.cuh"

namespace scout {

// Parameter layout (float offsets into the flat parameter vector).
constexpr int kOffW1 = 0;                               // [kMlpIn][kMlpHidden]
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHidden;    // [kMlpHidden]
constexpr int kOffW2 = kOffB1 + kMlpHidden;             // [kMlpHidden][kMlpHidden]
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;             // [kMlpHidden][kMlpOut]
constexpr int kOffB3 = kOffW3 + kMlpHidden * kMlpOut;
static_assert(kOffB3 + kMlpOut == kMlpParams, "parameter layout");

constexpr int kTrainBatch = 4096;   // samples per training step
constexpr int kGradSlice = 64;      // samples per wave and partial gradient sum
constexpr int kMaxSamples = 1 << 22;

struct MlpState {
  DevBuf<float> master, m, v, grad, partial;  // fp32 master weights, Adam moments
  DevBuf<uint16_t> w16;                        // bf16 weights used by every forward pass
  DevBuf<float> samples;                       // [kMaxSamples][kMlpIn + kMlpOut]
  DevBuf<float> feat_tmp;                      // features of this batch's sampled candidates
  DevBuf<int32_t> lvl0, subtree, maxlvl;       // per sampled candidate
  DevBuf<int32_t> root;                        // per vertex: candidate a cascade started from
  DevBuf<float> out_tmp;                       // predictions of this batch's sampled candidates
  DevBuf<float> io;                            // staging for the test entry points
  int64_t nsamples = 0;
  int64_t step = 0;
  bool initialised = false;
  bool logging = false;
  bool trained = false;
};

namespace {

constexpr int kRow = kMlpIn + kMlpOut;  // floats per stored sample

__host__ __device__ inline float bf16_to_f32(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// Round to nearest even (NaN stays NaN).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }
__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + expf(-z)); }

__device__ __forceinline__ float wave_sum_f(float v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// The wave64 mapping: lane j owns hidden unit j of both layers (kMlpHidden == kWave) and
// keeps its weights in registers for every sample the wave processes.
static_assert(kMlpHidden == kWave, "one hidden unit per lane");

struct LaneWeights {
  float w1[kMlpIn];       // W1[:, j]
  float b1;               // b1[j]
  float w2c[kMlpHidden];  // W2[:, j]  (forward)
  float b2;               // b2[j]
  float w3[kMlpOut];      // W3[j, :]
  float b3[kMlpOut];      // b3 (every lane)
};

__device__ __forceinline__ void load_lane_weights(const uint16_t* w16, LaneWeights& W) {
  int j = lane_id();
#pragma unroll
  for (int i = 0; i < kMlpIn; i++) W.w1[i] = bf16_to_f32(w16[kOffW1 + i * kMlpHidden + j]);
  W.b1 = bf16_to_f32(w16[kOffB1 + j]);
#pragma unroll
  for (int i = 0; i < kMlpHidden; i++) W.w2c[i] = bf16_to_f32(w16[kOffW2 + i * kMlpHidden + j]);
  W.b2 = bf16_to_f32(w16[kOffB2 + j]);
#pragma unroll
  for (int k = 0; k < kMlpOut; k++) {
    W.w3[k] = bf16_to_f32(w16[kOffW3 + j * kMlpOut + k]);
    W.b3[k] = bf16_to_f32(w16[kOffB3 + k]);
  }
}

// Forward pass of one sample by one wave (every lane holds x). Lane j receives h1[j], h2[j];
// every lane receives out[].
__device__ __forceinline__ void wave_forward(const LaneWeights& W, const float (&x)[kMlpIn], float& h1, float& h2,
                                             float (&out)[kMlpOut]) {
  float a = W.b1;
#pragma unroll
  for (int i = 0; i < kMlpIn; i++) a = fmaf(x[i], W.w1[i], a);
  h1 = relu(a);
  float b = W.b2;
#pragma unroll
  for (int i = 0; i < kMlpHidden; i++) b = fmaf(__shfl(h1, i), W.w2c[i], b);
  h2 = relu(b);
#pragma unroll
  for (int k = 0; k < kMlpOut; k++) out[k] = wave_sum_f(h2 * W.w3[k]) + W.b3[k];
}

// Loss of one sample: BCE(out0, y0) + 0.5 (out1 - y1)^2 + 0.5 (out2 - y2)^2, and its gradient
// with respect to the outputs.
__device__ __forceinline__ float loss_grad(const float (&out)[kMlpOut], const float* y, float (&d3)[kMlpOut]) {
  float z = out[0], t = y[0];
  // BCE with logits, stable form: max(z,0) - z t + log(1 + exp(-|z|))
  float l = fmaxf(z, 0.f) - z * t + log1pf(expf(-fabsf(z)));
  d3[0] = sigmoid(z) - t;
  float e1 = out[1] - y[1], e2 = out[2] - y[2];
  d3[1] = e1;
  d3[2] = e2;
  return l + 0.5f * (e1 * e1 + e2 * e2);
}

// Sample s of a training step: index into the store (a fixed stride walk = a permutation when
// the stride is coprime with the sample count).
__device__ __forceinline__ int64_t sample_index(int64_t start, int64_t s, int64_t stride, int64_t n) {
  return (int64_t)(((unsigned long long)(start + s) * (unsigned long long)stride) % (unsigned long long)n);
}

// One wave per slice of kGradSlice samples: forward, loss, backward, and the slice's summed
// gradient written to partial[slice][param] (the loss gradient is averaged over nb samples).
__global__ void __launch_bounds__(kWave) k_mlp_train_slice(const uint16_t* w16, const float* samples, int64_t nstore,
                                                          int64_t start, int64_t stride, int nb, float* partial,
                                                          float* loss_sum) {
  LaneWeights W;
  load_lane_weights(w16, W);
  int j = lane_id();
  // lane j also holds row j of W2 for the backward pass
  float w2r[kMlpHidden];
#pragma unroll
  for (int i = 0; i < kMlpHidden; i++) w2r[i] = bf16_to_f32(w16[kOffW2 + j * kMlpHidden + i]);
  float gw1[kMlpIn] = {}, gw2[kMlpHidden] = {}, gw3[kMlpOut] = {}, gb3[kMlpOut] = {};
  float gb1 = 0.f, gb2 = 0.f, loss = 0.f;
  int slice = blockIdx.x;
  int s0 = slice * kGradSlice, s1 = min(nb, s0 + kGradSlice);
  float inv = 1.f / (float)nb;
  for (int s = s0; s < s1; s++) {
    const float* row = samples + sample_index(start, s, stride, nstore) * kRow;
    float x[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) x[i] = row[i];
    float h1, h2, out[kMlpOut], d3[kMlpOut];
    wave_forward(W, x, h1, h2, out);
    loss += loss_grad(out, row + kMlpIn, d3);
#pragma unroll
    for (int k = 0; k < kMlpOut; k++) d3[k] *= inv;
    // d2[j] = (W3[j,:] . d3) relu'(z2[j]);  d1[j] = (W2[j,:] . d2) relu'(z1[j])
    float a = 0.f;
#pragma unroll
    for (int k = 0; k < kMlpOut; k++) a = fmaf(W.w3[k], d3[k], a);
    float d2 = h2 > 0.f ? a : 0.f;
    float b = 0.f;
#pragma unroll
    for (int i = 0; i < kMlpHidden; i++) b = fmaf(w2r[i], __shfl(d2, i), b);
    float d1 = h1 > 0.f ? b : 0.f;
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) gw1[i] = fmaf(x[i], d1, gw1[i]);
    gb1 += d1;
#pragma unroll
    for (int i = 0; i < kMlpHidden; i++) gw2[i] = fmaf(__shfl(h1, i), d2, gw2[i]);
    gb2 += d2;
#pragma unroll
    for (int k = 0; k < kMlpOut; k++) {
      gw3[k] = fmaf(h2, d3[k], gw3[k]);
      gb3[k] += d3[k];
    }
  }
  float* part = partial + (int64_t)slice * kMlpParams;
#pragma unroll
  for (int i = 0; i < kMlpIn; i++) part[kOffW1 + i * kMlpHidden + j] = gw1[i];
  part[kOffB1 + j] = gb1;
#pragma unroll
  for (int i = 0; i < kMlpHidden; i++) part[kOffW2 + i * kMlpHidden + j] = gw2[i];
  part[kOffB2 + j] = gb2;
#pragma unroll
  for (int k = 0; k < kMlpOut; k++) part[kOffW3 + j * kMlpOut + k] = gw3[k];
  if (j == 0) {
#pragma unroll
    for (int k = 0; k < kMlpOut; k++) part[kOffB3 + k] = gb3[k];
    atomicAdd(loss_sum, loss);
  }
}

// Thread per parameter: slices summed in order (deterministic), then Adam on the fp32 master
// weight and its bf16 copy. update = 0 only writes the gradient.
__global__ void k_mlp_reduce_adam(const float* partial, int nslices, float* grad, float* master, float* m, float* v,
                                  uint16_t* w16, float lr, float b1t, float b2t, int update) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  float g = 0.f;
  for (int k = 0; k < nslices; k++) g += partial[(int64_t)k * kMlpParams + p];
  grad[p] = g;
  if (!update) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  float mm = beta1 * m[p] + (1.f - beta1) * g;
  float vv = beta2 * v[p] + (1.f - beta2) * g * g;
  m[p] = mm;
  v[p] = vv;
  float mh = mm / (1.f - b1t), vh = vv / (1.f - b2t);
  float w = master[p] - lr * mh / (sqrtf(vh) + eps);
  master[p] = w;
  w16[p] = f32_to_bf16(w);
}

__global__ void k_mlp_round(const float* master, uint16_t* w16) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) w16[p] = f32_to_bf16(master[p]);
}

// Inference: wave per sample (grid-stride); out[n][kMlpOut]. With n_dev != nullptr the row
// count is min(*n_dev, n).
__global__ void __launch_bounds__(kBlock) k_mlp_forward(const uint16_t* w16, const float* feats, int64_t n,
                                                        const int64_t* n_dev, float* out) {
  LaneWeights W;
  load_lane_weights(w16, W);
  if (n_dev && *n_dev < n) n = *n_dev;
  for (int64_t s = global_wave(); s < n; s += total_waves()) {
    float x[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) x[i] = feats[s * kMlpIn + i];
    float h1, h2, o[kMlpOut];
    wave_forward(W, x, h1, h2, o);
    if (lane_id() == 0) {
#pragma unroll
      for (int k = 0; k < kMlpOut; k++) out[s * kMlpOut + k] = o[k];
    }
  }
}

// ---------------------------------------------------------------------------------------
// Features and outcomes of the sampled candidates of a batch
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ float lg(double x) { return (float)log1p(x); }

// Thread per sampled candidate c (the first ns of the list). Features, all from integer
// state before the repair:
//   0 log1p deg(c)            1 log1p deg(P[c])          2 L[c] / 16
//   3 log1p #neighbours at L[c]-1, 4 at L[c], 5 at L[c]+1   (first kShortScan entries)
//   6 scanned fraction of the row   7 (L[c] - cmin) / 16
//   8 log1p candidates / 16    9 log1p deletion lines / 16
//   10 [no neighbour at L[c]-1 in the scan] * log1p deg(c)   11 log1p n / 32
constexpr int kScan = 16;

// Counts come from the candidate pass's mailbox: box[0] candidates, box[1] smallest candidate
// level, box[3] deletion lines (no read-back needed before the features).
__global__ void k_mlp_features(GView g, const int32_t* cand, const int64_t* box, int64_t sample, const int32_t* level,
                               const int32_t* parent, float* feats, int32_t* lvl0, int32_t* subtree, int32_t* maxlvl) {
  int64_t ncand = box[0], ndel = box[3];
  int32_t cmin = (int32_t)box[1];
  int64_t ns = ncand < sample ? ncand : sample;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  if (subtree) {
    subtree[i] = 0;
    maxlvl[i] = -1;
  }
  int32_t c = cand[i];
  int32_t lc = level[c];
  Row r = row_of(g, c);
  int64_t len = row_len(r);
  int32_t p = parent[c];
  int64_t plen = p >= 0 ? row_len(row_of(g, p)) : 0;
  int up = 0, same = 0, down = 0;
  int64_t scan = len < kScan ? len : kScan;
  for (int64_t k = 0; k < scan; k++) {
    int32_t x = row_at(g, r, k);
    if (x < 0) continue;
    int32_t lx = level[x];
    up += lx == lc - 1;
    same += lx == lc;
    down += lx == lc + 1;
  }
  float* f = feats + i * kMlpIn;
  f[0] = lg((double)len);
  f[1] = lg((double)plen);
  f[2] = (float)lc / 16.f;
  f[3] = lg((double)up);
  f[4] = lg((double)same);
  f[5] = lg((double)down);
  f[6] = len > 0 ? (float)scan / (float)len : 1.f;
  f[7] = (float)(lc - cmin) / 16.f;
  f[8] = lg((double)ncand) / 16.f;
  f[9] = lg((double)ndel) / 16.f;
  f[10] = up == 0 ? f[0] : 0.f;
  f[11] = lg((double)g.n) / 32.f;
  lvl0[i] = lc;
}

// Outcomes after the repair, appended with the features to the sample store.
__global__ void k_mlp_log(const int32_t* cand, int64_t ns, const int32_t* level, const float* feats,
                          const int32_t* lvl0, const int32_t* subtree, const int32_t* maxlvl, float* samples,
                          int64_t base, int64_t cap) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns || base + i >= cap) return;
  float* row = samples + (base + i) * kRow;
  for (int k = 0; k < kMlpIn; k++) row[k] = feats[i * kMlpIn + k];
  int32_t c = cand[i];
  int32_t size = subtree[i];
  row[kMlpIn + 0] = level[c] < 0 ? 1.f : 0.f;
  row[kMlpIn + 1] = (float)log1p((double)size);
  row[kMlpIn + 2] = size > 0 ? (float)(maxlvl[i] - lvl0[i] + 1) : 0.f;
}

__global__ void k_pred_init(int64_t* box) {
  box[0] = 0;
  box[1] = -1;
  box[2] = INT64_MAX;
  box[3] = 0;
}

// Aggregates of the predictions: box[0] = sum of predicted cascade sizes (rounded), box[1] =
// the largest predicted last level L[c] + depth - 1, box[2] = minus the largest predicted depth
// (levels from a candidate to the deepest vertex it invalidates), box[3] = the expected number of vertices
// cut off vertex 0's component in sixteenths: per candidate P(c ends unreached) times its
// predicted cascade (c and the vertices it invalidates).
__global__ void k_mlp_aggregate(const float* out, const int32_t* lvl0, const int64_t* ncand, int64_t sample,
                                int64_t* box) {
  int64_t ns = *ncand < sample ? *ncand : sample;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t size = 0, last = -1, cut16 = 0, neg_depth = INT64_MAX;
  if (i < ns) {
    float s = expm1f(fmaxf(out[i * kMlpOut + 1], 0.f));
    float d = fmaxf(out[i * kMlpOut + 2], 0.f);
    float p = 1.f / (1.f + expf(-fminf(fmaxf(out[i * kMlpOut + 0], -30.f), 30.f)));
    size = (int64_t)fminf(s + 0.5f, 1e15f);
    last = lvl0[i] + (int64_t)ceilf(d) - 1;
    neg_depth = -(int64_t)ceilf(d);
    cut16 = (int64_t)fminf(16.f * p * (1.f + s) + 0.5f, 1e15f);
  }
  block_add(size, &box[0]);
  block_min_max(neg_depth, last, &box[2], &box[1]);
  block_add(cut16, &box[3]);
}

int ensure_mlp(Ctx& c) {
  if (!c.mlp) c.mlp = new MlpState();
  MlpState& s = *c.mlp;
  SCOUT_CHECK(s.master.reserve(kMlpParams));
  SCOUT_CHECK(s.m.reserve(kMlpParams));
  SCOUT_CHECK(s.v.reserve(kMlpParams));
  SCOUT_CHECK(s.grad.reserve(kMlpParams));
  SCOUT_CHECK(s.w16.reserve(kMlpParams));
  SCOUT_CHECK(s.partial.reserve((int64_t)(kTrainBatch / kGradSlice + 1) * kMlpParams));
  SCOUT_CHECK(s.feat_tmp.reserve((int64_t)kMlpSample * kMlpIn));
  SCOUT_CHECK(s.out_tmp.reserve((int64_t)kMlpSample * kMlpOut));
  SCOUT_CHECK(s.lvl0.reserve(kMlpSample));
  SCOUT_CHECK(s.subtree.reserve(kMlpSample));
  SCOUT_CHECK(s.maxlvl.reserve(kMlpSample));
  return SCOUT_OK;
}

// He-uniform initial weights from a counter-based RNG; biases 0.
void init_params(std::vector<float>& p, uint64_t seed) {
  p.assign(kMlpParams, 0.f);
  uint64_t key = stream_key(seed, 0x4d4c50);
  auto fill = [&](int off, int fan_in, int count) {
    float bound = std::sqrt(6.f / (float)fan_in);
    for (int i = 0; i < count; i++) {
      double u = (double)(draw(key, (uint64_t)(off + i)) >> 11) * (1.0 / 9007199254740992.0);
      p[(size_t)(off + i)] = (float)((2.0 * u - 1.0) * bound);
    }
  };
  fill(kOffW1, kMlpIn, kMlpIn * kMlpHidden);
  fill(kOffW2, kMlpHidden, kMlpHidden * kMlpHidden);
  fill(kOffW3, kMlpHidden, kMlpHidden * kMlpOut);
}

int set_params(Ctx& c, const float* host) {
  SCOUT_CHECK(ensure_mlp(c));
  MlpState& s = *c.mlp;
  SCOUT_TRY(hipMemcpyAsync(s.master.p, host, kMlpParams * sizeof(float), hipMemcpyHostToDevice, c.s));
  SCOUT_TRY(hipMemsetAsync(s.m.p, 0, kMlpParams * sizeof(float), c.s));
  SCOUT_TRY(hipMemsetAsync(s.v.p, 0, kMlpParams * sizeof(float), c.s));
  k_mlp_round<<<grid_for(kMlpParams), kBlock, 0, c.s>>>(s.master.p, s.w16.p);
  SCOUT_TRY(hipStreamSynchronize(c.s));
  s.step = 0;
  s.initialised = true;
  return SCOUT_OK;
}

// One forward/backward over nb samples of the store starting at walk position `start`; the
// gradient lands in s.grad, the summed loss in *loss. update = 1 also takes an Adam step.
int train_step(Ctx& c, const float* store, int64_t nstore, int64_t start, int64_t stride, int nb, float lr,
               int update, float* loss_dev) {
  MlpState& s = *c.mlp;
  int nslices = (nb + kGradSlice - 1) / kGradSlice;
  k_mlp_train_slice<<<nslices, kWave, 0, c.s>>>(s.w16.p, store, nstore, start, stride, nb, s.partial.p, loss_dev);
  if (update) s.step++;
  float b1t = (float)std::pow(0.9, (double)std::max<int64_t>(s.step, 1));
  float b2t = (float)std::pow(0.999, (double)std::max<int64_t>(s.step, 1));
  k_mlp_reduce_adam<<<grid_for(kMlpParams), kBlock, 0, c.s>>>(s.partial.p, nslices, s.grad.p, s.master.p, s.m.p,
                                                              s.v.p, s.w16.p, lr, b1t, b2t, update);
  return SCOUT_OK;
}

int64_t coprime_stride(int64_t n, uint64_t seed) {
  if (n <= 1) return 1;
  int64_t st = (int64_t)(splitmix64(seed) % (uint64_t)n) | 1;
  auto gcd = [](int64_t a, int64_t b) {
    while (b) {
      int64_t t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  while (gcd(st, n) != 1) st = (st + 2) % n;
  return st == 0 ? 1 : st;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Hooks used by the BFS repair
// ---------------------------------------------------------------------------------------

bool mlp_logging(Ctx& c) { return c.mlp && c.mlp->logging; }
bool mlp_ready(Ctx& c) { return c.mlp && c.mlp->trained; }

int mlp_begin_batch(Ctx& c, const int32_t* cand, const int64_t* d_box, const int32_t* L, const int32_t* P,
                    int64_t sample) {
  SCOUT_CHECK(ensure_mlp(c));
  MlpState& s = *c.mlp;
  bool log = s.logging;
  k_mlp_features<<<grid_for(sample), kBlock, 0, c.s>>>(c.g.view(), cand, d_box, sample, L, P, s.feat_tmp.p, s.lvl0.p,
                                                       log ? s.subtree.p : nullptr, log ? s.maxlvl.p : nullptr);
  return SCOUT_OK;
}

int32_t* mlp_root(Ctx& c) {
  if (!mlp_logging(c)) return nullptr;
  MlpState& s = *c.mlp;
  if (s.root.reserve((int64_t)c.g.n + 1) != SCOUT_OK) return nullptr;
  return s.root.p;
}
int32_t* mlp_subtree(Ctx& c) { return mlp_logging(c) ? c.mlp->subtree.p : nullptr; }
int32_t* mlp_maxlvl(Ctx& c) { return mlp_logging(c) ? c.mlp->maxlvl.p : nullptr; }

int mlp_predict_async(Ctx& c, const int64_t* d_ncand, int64_t sample, int64_t* d_out) {
  MlpState& s = *c.mlp;
  // 8 samples per wave: the lane weights are loaded once per wave
  k_mlp_forward<<<(unsigned)std::max<int64_t>(1, sample / 8 * kWave / kBlock), kBlock, 0, c.s>>>(
      s.w16.p, s.feat_tmp.p, sample, d_ncand, s.out_tmp.p);
  k_pred_init<<<1, 1, 0, c.s>>>(d_out);
  k_mlp_aggregate<<<grid_for(sample), kBlock, 0, c.s>>>(s.out_tmp.p, s.lvl0.p, d_ncand, sample, d_out);
  return SCOUT_OK;
}

int mlp_end_batch(Ctx& c, const int32_t* cand, int64_t ns, const int32_t* L) {
  MlpState& s = *c.mlp;
  if (!s.logging || ns == 0) return SCOUT_OK;
  SCOUT_CHECK(s.samples.reserve((int64_t)kMaxSamples * kRow));
  k_mlp_log<<<grid_for(ns), kBlock, 0, c.s>>>(cand, ns, L, s.feat_tmp.p, s.lvl0.p, s.subtree.p, s.maxlvl.p,
                                              s.samples.p, s.nsamples, kMaxSamples);
  s.nsamples = std::min<int64_t>(s.nsamples + ns, kMaxSamples);
  return SCOUT_OK;
}

void mlp_release(Ctx& c) {
  if (!c.mlp) return;
  MlpState& s = *c.mlp;
  DevBuf<float>* fs[] = {&s.master, &s.m, &s.v, &s.grad, &s.partial, &s.samples, &s.feat_tmp, &s.out_tmp,
                         &s.io};
  for (auto* b : fs) b->release();
  s.w16.release();
  s.lvl0.release();
  s.subtree.release();
  s.maxlvl.release();
  s.root.release();
  delete c.mlp;
  c.mlp = nullptr;
}

}  // namespace scout

using namespace scout;

extern "C" {

int scout_mlp_init(scout_ctx* x, uint64_t seed) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  std::vector<float> p;
  init_params(p, seed);
  SCOUT_CHECK(set_params(c, p.data()));
  c.mlp->trained = false;
  return SCOUT_OK;
}

int scout_mlp_set_params(scout_ctx* x, const float* params) {
  if (!x || !params) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  SCOUT_CHECK(set_params(c, params));
  c.mlp->trained = true;
  return SCOUT_OK;
}

int scout_mlp_get_params(scout_ctx* x, float* params, int rounded) {
  if (!x || !params) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  if (!c.mlp || !c.mlp->initialised) {
    set_error("mlp: not initialised");
    return SCOUT_ERR_STATE;
  }
  MlpState& s = *c.mlp;
  if (rounded) {
    std::vector<uint16_t> h(kMlpParams);
    SCOUT_TRY(hipMemcpyAsync(h.data(), s.w16.p, kMlpParams * 2, hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
    for (int i = 0; i < kMlpParams; i++) params[i] = bf16_to_f32(h[(size_t)i]);
  } else {
    SCOUT_TRY(hipMemcpyAsync(params, s.master.p, kMlpParams * 4, hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
  }
  return SCOUT_OK;
}

int scout_mlp_forward(scout_ctx* x, const float* feats, int64_t n, float* out) {
  if (!x || n < 0 || (n > 0 && (!feats || !out))) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  if (!c.mlp || !c.mlp->initialised) {
    set_error("mlp: not initialised");
    return SCOUT_ERR_STATE;
  }
  if (n == 0) return SCOUT_OK;
  MlpState& s = *c.mlp;
  SCOUT_CHECK(s.io.reserve(n * (kMlpIn + kMlpOut)));
  float* df = s.io.p;
  float* dout = s.io.p + n * kMlpIn;
  SCOUT_TRY(hipMemcpyAsync(df, feats, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, c.s));
  k_mlp_forward<<<std::min<unsigned>(grid_for(n * kWave), kListGrid), kBlock, 0, c.s>>>(s.w16.p, df, n, nullptr,
                                                                                          dout);
  SCOUT_TRY(hipMemcpyAsync(out, dout, (size_t)n * kMlpOut * 4, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  return SCOUT_OK;
}

int scout_mlp_grad(scout_ctx* x, const float* feats, const float* targets, int64_t n, float* grad, float* loss) {
  if (!x || !feats || !targets || !grad || n <= 0 || n > kTrainBatch) {
    set_error("mlp_grad: 1 <= n <= 4096 samples");
    return SCOUT_ERR_ARG;
  }
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  if (!c.mlp || !c.mlp->initialised) {
    set_error("mlp: not initialised");
    return SCOUT_ERR_STATE;
  }
  MlpState& s = *c.mlp;
  std::vector<float> rows((size_t)n * kRow);
  for (int64_t i = 0; i < n; i++) {
    for (int k = 0; k < kMlpIn; k++) rows[(size_t)(i * kRow + k)] = feats[i * kMlpIn + k];
    for (int k = 0; k < kMlpOut; k++) rows[(size_t)(i * kRow + kMlpIn + k)] = targets[i * kMlpOut + k];
  }
  SCOUT_CHECK(s.io.reserve(n * kRow + 1));
  SCOUT_TRY(hipMemcpyAsync(s.io.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, c.s));
  float* dloss = s.io.p + n * kRow;
  SCOUT_TRY(hipMemsetAsync(dloss, 0, 4, c.s));
  SCOUT_CHECK(train_step(c, s.io.p, n, 0, 1, (int)n, 0.f, 0, dloss));
  SCOUT_TRY(hipMemcpyAsync(grad, s.grad.p, kMlpParams * 4, hipMemcpyDeviceToHost, c.s));
  float l = 0.f;
  SCOUT_TRY(hipMemcpyAsync(&l, dloss, 4, hipMemcpyDeviceToHost, c.s));
  SCOUT_TRY(hipStreamSynchronize(c.s));
  if (loss) *loss = l / (float)n;
  return SCOUT_OK;
}

int scout_mlp_log(scout_ctx* x, int on) {
  if (!x) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  SCOUT_CHECK(ensure_mlp(c));
  c.mlp->logging = on != 0;
  return SCOUT_OK;
}

int scout_mlp_samples(scout_ctx* x, float* rows, int64_t cap, int64_t* count) {
  if (!x || !count) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  int64_t n = c.mlp ? c.mlp->nsamples : 0;
  *count = n;
  if (rows && n > 0) {
    int64_t k = std::min(n, cap);
    SCOUT_TRY(hipMemcpyAsync(rows, c.mlp->samples.p, (size_t)k * kRow * 4, hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
  }
  return SCOUT_OK;
}

int scout_mlp_train(scout_ctx* x, int epochs, float lr, uint64_t seed, float* first_loss, float* last_loss) {
  if (!x || epochs < 1) return SCOUT_ERR_ARG;
  Ctx& c = x->c;
  SCOUT_TRY(hipSetDevice(c.dev));
  if (!c.mlp || !c.mlp->initialised) {
    set_error("mlp: scout_mlp_init first");
    return SCOUT_ERR_STATE;
  }
  MlpState& s = *c.mlp;
  int64_t n = s.nsamples;
  if (n < kTrainBatch) {
    set_error("mlp_train: fewer logged samples than one training batch (4096)");
    return SCOUT_ERR_STATE;
  }
  SCOUT_CHECK(s.io.reserve(2));
  float* dloss = s.io.p;
  int64_t stride = coprime_stride(n, seed);
  int64_t steps = n / kTrainBatch;
  for (int e = 0; e < epochs; e++) {
    SCOUT_TRY(hipMemsetAsync(dloss, 0, 4, c.s));
    for (int64_t k = 0; k < steps; k++)
      SCOUT_CHECK(train_step(c, s.samples.p, n, k * kTrainBatch, stride, kTrainBatch, lr, 1, dloss));
    float l = 0.f;
    SCOUT_TRY(hipMemcpyAsync(&l, dloss, 4, hipMemcpyDeviceToHost, c.s));
    SCOUT_TRY(hipStreamSynchronize(c.s));
    l /= (float)(steps * kTrainBatch);
    if (e == 0 && first_loss) *first_loss = l;
    if (last_loss) *last_loss = l;
  }
  s.trained = true;
  return SCOUT_OK;
}

}  // extern "C"
