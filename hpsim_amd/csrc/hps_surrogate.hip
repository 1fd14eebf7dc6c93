// hps_surrogate.hip -- the SCHEFF field surrogate network: parameters,
// forward and backward passes, Adam training, on the hps_nn kernels.
//
// Architecture (a dilated convolutional operator on the field mesh, y = zp in
// [0, nz], x = rp in [0, nr]):
//   in (4 channels: normalised charge, log mesh aspect, z, r)
//   L0  conv3x3 4 -> C, ReLU
//   L1..L5  x += ReLU(conv3x3 C -> C, dilation 2, 4, 8, 16, 32)
//   L6  conv1x1 C -> 16 (channels 0, 1: E_r, E_z in mesh units)
// The receptive field (127 rows) spans the mesh, as the Coulomb field of
// every charge reaches every field point.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "hpsim_amd.h"
#include "hps_nn.hpp"
#include "hps_surrogate.hpp"

namespace hps {

int fail_msg(const std::string& m);   // hps_engine.hip

#define NN_CHECK(expr)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail_msg(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

struct Layer {
  int Cin, Cout, taps, dil, relu, resid;
  size_t w_off, b_off;   // offsets into the flat parameter vector
};

}  // namespace

struct NetImpl {
  int H = 0, W = 0, width = 0, max_batch = 0;
  std::vector<Layer> layers;
  size_t nparam = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wt = nullptr, *wb = nullptr;   // per-layer BF16 copies, at w_off
  std::vector<float*> out, act;            // per layer, [max_batch][H][W][Cout]
  float *x = nullptr, *tgt = nullptr, *dpred = nullptr, *loss = nullptr;
  float *d0 = nullptr, *d1 = nullptr;      // backward activations grads [max_batch][H][W][C]
  float *data_x = nullptr, *data_y = nullptr;
  int ndata = 0;
  int32_t* idx = nullptr;
  int step = 0;
  float out_scale = 1.0f;                  // the field is the output times this
  hipStream_t s = nullptr;
};

namespace {

constexpr int kInC = 4, kOutC = 16;
constexpr int kDil[] = {1, 2, 4, 8, 16, 32};

void build_layers(NetImpl* n)
{
  size_t off = 0;
  auto add = [&](int cin, int cout, int taps, int dil, int relu, int resid) {
    Layer l{cin, cout, taps, dil, relu, resid, off, off + (size_t)cout * taps * cin};
    off = l.b_off + cout;
    n->layers.push_back(l);
  };
  add(kInC, n->width, 9, 1, 1, 0);
  for (int k = 1; k < 6; ++k) add(n->width, n->width, 9, kDil[k], 1, 1);
  add(n->width, kOutC, 1, 1, 0, 0);
  n->nparam = off;
}

void free_net(NetImpl* n)
{
  for (float* p : {n->w, n->g, n->m, n->v, n->x, n->tgt, n->dpred, n->loss, n->d0, n->d1,
                   n->data_x, n->data_y})
    (void)hipFree(p);
  (void)hipFree(n->wt); (void)hipFree(n->wb); (void)hipFree(n->idx);
  for (float* p : n->out) (void)hipFree(p);
  for (float* p : n->act) (void)hipFree(p);
}

int refresh_bf16(NetImpl* n)
{
  for (const Layer& l : n->layers)
    NN_CHECK(nn::weights_bf16(n->w + l.w_off, n->wt + l.w_off, n->wb + l.w_off, l.Cout, l.taps,
                              l.Cin, n->s));
  return 0;
}

nn::ConvGeom geom(const NetImpl* n, const Layer& l, int B)
{
  return nn::ConvGeom{B, n->H, n->W, l.Cin, l.Cout, l.dil};
}

}  // namespace

// Forward pass of B samples from `in` ([B][H][W][4]); the prediction is
// layers.back()'s output ([B][H][W][16], channels 0, 1 used).
int net_forward(NetImpl* n, const float* in, int B, hipStream_t s)
{
  const float* x = in;
  for (size_t k = 0; k < n->layers.size(); ++k) {
    const Layer& l = n->layers[k];
    NN_CHECK(nn::conv_fwd(x, n->wt + l.w_off, n->w + l.b_off, l.resid ? x : nullptr,
                          l.resid ? n->act[k] : nullptr, n->out[k], geom(n, l, B), l.taps,
                          l.relu, s));
    x = n->out[k];
  }
  return 0;
}

const float* net_output(const NetImpl* n) { return n->out.back(); }
int net_out_channels() { return kOutC; }

// Gradients (accumulated into n->g, which the caller zeroes) of the mean
// relative-L2 loss of the batch in n->x / n->tgt; per-sample losses in n->loss.
static int net_backward(NetImpl* n, int B)
{
  const int HW = n->H * n->W;
  NN_CHECK(nn::rel_l2_loss(n->out.back(), kOutC, n->tgt, n->dpred, n->loss, HW, B, n->s));
  const float* dout = n->dpred;
  float* bufs[2] = {n->d0, n->d1};
  int nb = 0;
  for (int k = (int)n->layers.size() - 1; k >= 0; --k) {
    const Layer& l = n->layers[k];
    const float* in = k == 0 ? n->x : n->out[k - 1];
    // the ReLU mask: the activation before the residual add
    const float* act = l.resid ? n->act[k] : n->out[k];
    NN_CHECK(nn::conv_bwd_w(dout, act, in, n->g + l.w_off, n->g + l.b_off, geom(n, l, B), l.taps,
                            l.relu, n->s));
    if (k == 0) break;
    float* din = bufs[nb];
    NN_CHECK(nn::conv_bwd_data(dout, act, n->wb + l.w_off, l.resid ? dout : nullptr, din,
                               geom(n, l, B), l.taps, l.relu, n->s));
    dout = din;
    nb ^= 1;
  }
  return 0;
}

}  // namespace hps

using namespace hps;

struct hps_nn {
  NetImpl impl;
};

extern "C" {

int hps_nn_create(hps_nn** out, uint32_t H, uint32_t W, uint32_t width, uint32_t max_batch,
                  uint32_t seed)
{
  if (!out || H == 0 || W == 0 || max_batch == 0) return fail_msg("bad network shape");
  if (width % 16 != 0 || width == 0 || width > (uint32_t)nn::kMaxC)
    return fail_msg("network width must be a multiple of 16, at most 64");
  hps_nn* net = new hps_nn();
  NetImpl* n = &net->impl;
  n->H = (int)H; n->W = (int)W; n->width = (int)width; n->max_batch = (int)max_batch;
  build_layers(n);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { delete net; return fail_msg("no HIP device"); }
  bool ok = hipStreamCreate(&n->s) == hipSuccess;
  const size_t np = n->nparam * sizeof(float);
  for (float** p : {&n->w, &n->g, &n->m, &n->v}) ok = ok && hipMalloc(p, np) == hipSuccess;
  ok = ok && hipMalloc(&n->wt, n->nparam * 2) == hipSuccess &&
       hipMalloc(&n->wb, n->nparam * 2) == hipSuccess;
  const size_t pos = (size_t)max_batch * H * W;
  n->out.assign(n->layers.size(), nullptr);
  n->act.assign(n->layers.size(), nullptr);
  for (size_t k = 0; k < n->layers.size() && ok; ++k) {
    const size_t sz = pos * n->layers[k].Cout * sizeof(float);
    ok = hipMalloc(&n->out[k], sz) == hipSuccess;
    if (ok && n->layers[k].resid) ok = hipMalloc(&n->act[k], sz) == hipSuccess;
  }
  ok = ok && hipMalloc(&n->x, pos * kInC * 4) == hipSuccess &&
       hipMalloc(&n->tgt, pos * 2 * 4) == hipSuccess &&
       hipMalloc(&n->dpred, pos * kOutC * 4) == hipSuccess &&
       hipMalloc(&n->loss, (size_t)max_batch * 4) == hipSuccess &&
       hipMalloc(&n->d0, pos * nn::kMaxC * 4) == hipSuccess &&
       hipMalloc(&n->d1, pos * nn::kMaxC * 4) == hipSuccess &&
       hipMalloc(&n->idx, (size_t)max_batch * 4) == hipSuccess;
  if (!ok) { free_net(n); delete net; return fail_msg("device allocation failed for the network"); }
  // He-normal initialisation (std::mt19937_64 with the given seed), zero biases
  std::vector<float> w(n->nparam, 0.0f);
  std::mt19937_64 rng(seed);
  for (const Layer& l : n->layers) {
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)(l.taps * l.Cin)));
    const float scale = l.resid ? 0.25f : 1.0f;   // residual branches start small
    for (size_t i = 0; i < (size_t)l.Cout * l.taps * l.Cin; ++i) w[l.w_off + i] = scale * nd(rng);
  }
  if (hipMemcpy(n->w, w.data(), np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, np) != hipSuccess || hipMemset(n->v, 0, np) != hipSuccess ||
      refresh_bf16(n) || hipStreamSynchronize(n->s) != hipSuccess) {
    free_net(n);
    delete net;
    return fail_msg("network initialisation failed");
  }
  *out = net;
  return 0;
}

int hps_nn_destroy(hps_nn* net)
{
  if (!net) return 0;
  (void)hipStreamSynchronize(net->impl.s);
  free_net(&net->impl);
  (void)hipStreamDestroy(net->impl.s);
  delete net;
  return 0;
}

uint32_t hps_nn_num_params(const hps_nn* net) { return net ? (uint32_t)net->impl.nparam : 0; }

int hps_nn_get_params(hps_nn* net, float* host)
{
  if (!net || !host) return fail_msg("null argument");
  NetImpl* n = &net->impl;
  NN_CHECK(hipStreamSynchronize(n->s));
  NN_CHECK(hipMemcpy(host, n->w, n->nparam * 4, hipMemcpyDeviceToHost));
  return 0;
}

int hps_nn_set_params(hps_nn* net, const float* host)
{
  if (!net || !host) return fail_msg("null argument");
  NetImpl* n = &net->impl;
  NN_CHECK(hipStreamSynchronize(n->s));
  NN_CHECK(hipMemcpy(n->w, host, n->nparam * 4, hipMemcpyHostToDevice));
  if (refresh_bf16(n)) return -1;
  NN_CHECK(hipStreamSynchronize(n->s));
  return 0;
}

int hps_nn_set_data(hps_nn* net, const float* x, const float* y, uint32_t count)
{
  if (!net || !x || !y || count == 0) return fail_msg("bad training data");
  NetImpl* n = &net->impl;
  NN_CHECK(hipStreamSynchronize(n->s));
  (void)hipFree(n->data_x); (void)hipFree(n->data_y);
  n->data_x = n->data_y = nullptr;
  const size_t hw = (size_t)n->H * n->W;
  NN_CHECK(hipMalloc(&n->data_x, count * hw * kInC * 4));
  NN_CHECK(hipMalloc(&n->data_y, count * hw * 2 * 4));
  NN_CHECK(hipMemcpy(n->data_x, x, count * hw * kInC * 4, hipMemcpyHostToDevice));
  NN_CHECK(hipMemcpy(n->data_y, y, count * hw * 2 * 4, hipMemcpyHostToDevice));
  n->ndata = (int)count;
  return 0;
}

// steps of Adam on random batches of the data set; losses[k] = mean batch
// loss of step k (optional)
int hps_nn_train(hps_nn* net, uint32_t steps, uint32_t batch, float lr, uint32_t seed,
                 float* losses)
{
  if (!net) return fail_msg("null network");
  NetImpl* n = &net->impl;
  if (n->ndata == 0) return fail_msg("no training data");
  if (batch == 0 || (int)batch > n->max_batch) return fail_msg("batch exceeds the network's max");
  const size_t hw = (size_t)n->H * n->W;
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> pick(0, n->ndata - 1);
  std::vector<int32_t> idx(batch);
  std::vector<float> lb(batch);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t k = 0; k < steps; ++k) {
    for (auto& i : idx) i = pick(rng);
    NN_CHECK(hipMemcpyAsync(n->idx, idx.data(), batch * 4, hipMemcpyHostToDevice, n->s));
    NN_CHECK(nn::gather_rows(n->data_x, n->idx, n->x, (int)batch, (int)(hw * kInC), n->s));
    NN_CHECK(nn::gather_rows(n->data_y, n->idx, n->tgt, (int)batch, (int)(hw * 2), n->s));
    NN_CHECK(hipMemsetAsync(n->g, 0, n->nparam * 4, n->s));
    if (net_forward(n, n->x, (int)batch, n->s) || net_backward(n, (int)batch)) return -1;
    n->step += 1;
    const float lr_t = lr * std::sqrt(1.0f - std::pow(b2, (float)n->step)) /
                       (1.0f - std::pow(b1, (float)n->step));
    NN_CHECK(nn::adam(n->w, n->g, n->m, n->v, (int)n->nparam, lr_t, b1, b2, eps, n->s));
    if (refresh_bf16(n)) return -1;
    if (losses) {
      NN_CHECK(hipMemcpyAsync(lb.data(), n->loss, batch * 4, hipMemcpyDeviceToHost, n->s));
      NN_CHECK(hipStreamSynchronize(n->s));
      double sacc = 0;
      for (float v : lb) sacc += v;
      losses[k] = (float)(sacc / batch);
    }
  }
  NN_CHECK(hipStreamSynchronize(n->s));
  return 0;
}

// prediction (channels 0, 1, times the output scale) of `count` samples:
// x [count][H][W][4] -> out [count][H][W][2]
int hps_nn_predict(hps_nn* net, const float* x, uint32_t count, float* out)
{
  if (!net || !x || !out) return fail_msg("null argument");
  NetImpl* n = &net->impl;
  const size_t hw = (size_t)n->H * n->W;
  std::vector<float> full(hw * kOutC * n->max_batch);
  for (uint32_t c0 = 0; c0 < count; c0 += n->max_batch) {
    const int B = (int)std::min<uint32_t>(n->max_batch, count - c0);
    NN_CHECK(hipMemcpyAsync(n->x, x + c0 * hw * kInC, B * hw * kInC * 4, hipMemcpyHostToDevice, n->s));
    if (net_forward(n, n->x, B, n->s)) return -1;
    NN_CHECK(hipMemcpyAsync(full.data(), n->out.back(), B * hw * kOutC * 4, hipMemcpyDeviceToHost, n->s));
    NN_CHECK(hipStreamSynchronize(n->s));
    for (size_t i = 0; i < B * hw; ++i) {
      out[(c0 * hw + i) * 2] = full[i * kOutC] * n->out_scale;
      out[(c0 * hw + i) * 2 + 1] = full[i * kOutC + 1] * n->out_scale;
    }
  }
  return 0;
}

// loss of each sample and the gradient of their mean over all parameters
int hps_nn_loss_grad(hps_nn* net, const float* x, const float* y, uint32_t count, float* loss,
                     float* grad)
{
  if (!net || !x || !y || !loss || !grad) return fail_msg("null argument");
  NetImpl* n = &net->impl;
  if ((int)count > n->max_batch || count == 0) return fail_msg("count exceeds the network's max batch");
  const size_t hw = (size_t)n->H * n->W;
  NN_CHECK(hipMemcpyAsync(n->x, x, count * hw * kInC * 4, hipMemcpyHostToDevice, n->s));
  NN_CHECK(hipMemcpyAsync(n->tgt, y, count * hw * 2 * 4, hipMemcpyHostToDevice, n->s));
  NN_CHECK(hipMemsetAsync(n->g, 0, n->nparam * 4, n->s));
  if (net_forward(n, n->x, (int)count, n->s) || net_backward(n, (int)count)) return -1;
  NN_CHECK(hipMemcpyAsync(loss, n->loss, count * 4, hipMemcpyDeviceToHost, n->s));
  NN_CHECK(hipMemcpyAsync(grad, n->g, n->nparam * 4, hipMemcpyDeviceToHost, n->s));
  NN_CHECK(hipStreamSynchronize(n->s));
  return 0;
}

int hps_nn_set_output_scale(hps_nn* net, float v)
{
  if (!net || !(v > 0.0f)) return fail_msg("the output scale must be positive");
  net->impl.out_scale = v;
  return 0;
}

int hps_nn_get_output_scale(const hps_nn* net, float* v)
{
  if (!net || !v) return fail_msg("null argument");
  *v = net->impl.out_scale;
  return 0;
}

// layer table: for layer k, {Cin, Cout, taps, dilation, relu, residual, weight offset, bias offset}
int hps_nn_layers(const hps_nn* net, int32_t* table, uint32_t max_layers, uint32_t* nlayers)
{
  if (!net || !nlayers) return fail_msg("null argument");
  const NetImpl* n = &net->impl;
  *nlayers = (uint32_t)n->layers.size();
  for (uint32_t k = 0; table && k < max_layers && k < n->layers.size(); ++k) {
    const Layer& l = n->layers[k];
    int32_t* r = table + 8 * k;
    r[0] = l.Cin; r[1] = l.Cout; r[2] = l.taps; r[3] = l.dil; r[4] = l.relu; r[5] = l.resid;
    r[6] = (int32_t)l.w_off; r[7] = (int32_t)l.b_off;
  }
  return 0;
}

}  // extern "C"
