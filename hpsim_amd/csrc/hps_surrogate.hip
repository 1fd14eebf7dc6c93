// hps_surrogate.hip -- the SCHEFF field surrogate network object: layer
// table, parameters, forward and backward passes on the hps_unet kernels,
// Adam training on a device-resident dataset, and the C ABI (hps_nn_*).
//
// Architecture (Z0 = nz + 8 positions along z, level widths C0..C3):
//   lift   point  nr -> C0            enc0  same  C0 -> C0
//   down0  down   C0 -> C1            enc1  same  C1 -> C1
//   down1  down   C1 -> C2            enc2  same  C2 -> C2
//   down2  down   C2 -> C3            bottom: nbot x same C3 -> C3, dilation 1, 2, 4, ...
//   up2    up     C3 -> C2 + enc2     dec2  same  C2 -> C2
//   up1    up     C2 -> C1 + enc1     dec1  same  C1 -> C1
//   up0    up     C1 -> C0 + enc0     dec0  same  C0 -> C0
//   proj   point  C0 -> 2 (nr + 1), padded to a multiple of 16
// every layer with FiLM scales 1 + A1 v + A2 v^2 + A3 v^3 per channel, v =
// 2 tanh(u / 2) of the standardised log aspect u.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "hpsim_amd.h"
#include "hps_unet.hpp"

namespace hps {

int fail_msg(const std::string& m);   // hps_engine.hip

#define NN_CHECK(expr)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail_msg(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

struct Layer {
  int mode, cin, cout, taps, dil, skip;
  size_t w_off, a_off;
  int zin, zout;
};

constexpr int kPad = 8;   // zero positions after the nz bins

__global__ void k_add(float* __restrict__ dst, const float* __restrict__ src, size_t n)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

// u[i] = (lnk[idx[i]] - mu) * inv_sigma (idx null: identity)
__global__ void k_norm_u(const float* __restrict__ lnk, const uint32_t* __restrict__ idx,
                         float* __restrict__ u, int n, float mu, float inv_sigma)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) u[i] = (lnk[idx ? idx[i] : i] - mu) * inv_sigma;
}

__global__ void k_mean(const float* __restrict__ l, int n, float* __restrict__ out)
{
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += l[i];
    out[0] = s / (float)n;
  }
}

}  // namespace

struct NetImpl {
  int nr = 0, nz = 0, Z0 = 0, cin = 0, cout = 0, cpad = 0, zn = 0;
  std::vector<Layer> layers;
  size_t nparam = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wt = nullptr, *wb = nullptr;
  unet::FilmLayer* d_film = nullptr;       // scale offsets for max_b samples
  unet::FilmLayer* d_film_inf = nullptr;   // ... for inf_b samples
  size_t scale_per_sample = 0;       // sum of the layers' Cout
  // training / evaluation buffers for max_b samples
  int max_b = 0;
  std::vector<float*> pre, out, dy;
  float *scale = nullptr, *dpre = nullptr, *xb = nullptr, *tb = nullptr, *ub = nullptr;
  float *lossb = nullptr, *hist = nullptr;
  uint32_t* idx = nullptr;
  int hist_cap = 0;
  // inference buffers (the batch engine) for inf_b samples
  int inf_b = 0;
  std::vector<float*> iout;
  float* iscale = nullptr;
  // dataset
  NetData data{};
  float out_scale = 1.f, u_mu = 0.f, u_sigma = 1.f;
  int step = 0;
  hipStream_t s = nullptr;
};

namespace {

void build_layers(NetImpl* n, const uint32_t* wd, uint32_t nbot)
{
  const int C[4] = {(int)wd[0], (int)wd[1], (int)wd[2], (int)wd[3]};
  size_t off = 0;
  int z = n->Z0;
  auto add = [&](int mode, int cin, int cout, int dil, int skip) -> int {
    const int taps = mode == unet::kSame ? 3 : (mode == unet::kPoint ? 1 : 2);
    Layer l{mode, cin, cout, taps, dil, skip, off, off + (size_t)cout * taps * cin, z, 0};
    l.zout = mode == unet::kDown ? z / 2 : (mode == unet::kUp ? 2 * z : z);
    z = l.zout;
    off = l.a_off + (size_t)cout * 3;
    n->layers.push_back(l);
    return (int)n->layers.size() - 1;
  };
  add(unet::kPoint, n->cin, C[0], 1, -1);
  const int e0 = add(unet::kSame, C[0], C[0], 1, -1);
  add(unet::kDown, C[0], C[1], 1, -1);
  const int e1 = add(unet::kSame, C[1], C[1], 1, -1);
  add(unet::kDown, C[1], C[2], 1, -1);
  const int e2 = add(unet::kSame, C[2], C[2], 1, -1);
  add(unet::kDown, C[2], C[3], 1, -1);
  for (uint32_t k = 0; k < nbot; ++k) add(unet::kSame, C[3], C[3], 1 << k, -1);
  add(unet::kUp, C[3], C[2], 1, e2);
  add(unet::kSame, C[2], C[2], 1, -1);
  add(unet::kUp, C[2], C[1], 1, e1);
  add(unet::kSame, C[1], C[1], 1, -1);
  add(unet::kUp, C[1], C[0], 1, e0);
  add(unet::kSame, C[0], C[0], 1, -1);
  add(unet::kPoint, C[0], n->cpad, 1, -1);
  n->nparam = off;
}

void free_bufs(std::vector<float*>& v)
{
  for (float* p : v) (void)hipFree(p);
  v.clear();
}

void free_net(NetImpl* n)
{
  for (float* p : {n->w, n->g, n->m, n->v, n->scale, n->dpre, n->xb, n->tb, n->ub, n->lossb,
                   n->hist, n->iscale, n->data.x, n->data.y, n->data.u})
    (void)hipFree(p);
  (void)hipFree(n->wt); (void)hipFree(n->wb); (void)hipFree(n->idx); (void)hipFree(n->d_film);
  (void)hipFree(n->d_film_inf);
  (void)hipFree(n->data.count);
  free_bufs(n->pre); free_bufs(n->out); free_bufs(n->dy); free_bufs(n->iout);
}

int refresh_bf16(NetImpl* n)
{
  for (const Layer& l : n->layers)
    NN_CHECK(unet::bf16_copies(n->w + l.w_off, n->wt + l.w_off, n->wb + l.w_off, l.cout, l.taps,
                               l.cin, n->s));
  return 0;
}

size_t act_size(const Layer& l, int b) { return (size_t)b * l.zout * l.cout; }

// Forward pass over B samples (or the device count dyn_b <= B): FiLM scales
// into `scale`, layer outputs into out (pre-FiLM products into pre when
// given, for the backward pass).
int forward(NetImpl* n, const float* x, const float* u, const uint32_t* dyn_b, int B,
            const unet::FilmLayer* film, std::vector<float*>& out, float* scale,
            std::vector<float*>* pre, hipStream_t s)
{
  NN_CHECK(unet::film(n->w, film, (int)n->layers.size(), u, dyn_b, B, scale, s));
  const float* h = x;
  size_t soff = 0;
  for (size_t k = 0; k < n->layers.size(); ++k) {
    const Layer& l = n->layers[k];
    unet::ConvArgs a{};
    a.in = h;
    a.w = n->wt + l.w_off;
    a.pre = pre ? (*pre)[k] : nullptr;
    a.out = out[k];
    a.scale = scale + soff * B;
    a.skip = l.skip >= 0 ? out[l.skip] : nullptr;
    a.dyn_b = dyn_b;
    a.B = B; a.Zin = l.zin; a.Zout = l.zout; a.Cin = l.cin; a.Cout = l.cout;
    a.taps = l.taps; a.dil = l.dil; a.mode = l.mode;
    NN_CHECK(unet::conv(a, s));
    soff += l.cout;
    h = out[k];
  }
  return 0;
}

// The FiLM table for B samples per layer (scale offsets soff * B).
int upload_film(NetImpl* n, int B, unet::FilmLayer* dst)
{
  std::vector<unet::FilmLayer> fl(n->layers.size());
  size_t soff = 0;
  for (size_t k = 0; k < n->layers.size(); ++k) {
    fl[k].cout = n->layers[k].cout;
    fl[k].a_off = n->layers[k].a_off;
    fl[k].s_off = soff * B;
    soff += n->layers[k].cout;
  }
  NN_CHECK(hipMemcpyAsync(dst, fl.data(), fl.size() * sizeof(unet::FilmLayer),
                          hipMemcpyHostToDevice, n->s));
  NN_CHECK(hipStreamSynchronize(n->s));
  return 0;
}

int alloc_train(NetImpl* n, int B)
{
  if (B <= n->max_b) return 0;
  free_bufs(n->pre); free_bufs(n->out); free_bufs(n->dy);
  for (float* p : {n->scale, n->dpre, n->xb, n->tb, n->ub, n->lossb}) (void)hipFree(p);
  n->max_b = B;
  size_t big = 0;
  for (const Layer& l : n->layers) {
    float *a = nullptr, *b = nullptr, *c = nullptr;
    NN_CHECK(hipMalloc(&a, act_size(l, B) * 4));
    NN_CHECK(hipMalloc(&b, act_size(l, B) * 4));
    NN_CHECK(hipMalloc(&c, act_size(l, B) * 4));
    n->pre.push_back(a); n->out.push_back(b); n->dy.push_back(c);
    big = std::max(big, act_size(l, B));
  }
  NN_CHECK(hipMalloc(&n->scale, n->scale_per_sample * B * 4));
  NN_CHECK(hipMalloc(&n->dpre, big * 4));
  NN_CHECK(hipMalloc(&n->xb, (size_t)B * n->Z0 * n->cin * 4));
  NN_CHECK(hipMemset(n->xb, 0, (size_t)B * n->Z0 * n->cin * 4));
  NN_CHECK(hipMalloc(&n->tb, (size_t)B * n->zn * n->cout * 4));
  NN_CHECK(hipMalloc(&n->ub, (size_t)B * 4));
  NN_CHECK(hipMalloc(&n->lossb, (size_t)B * 4));
  return 0;
}

// Loss of B samples already in xb, tb, ub; with grad, the gradient of the
// mean loss lands in n->g (zeroed here).
int loss_grad(NetImpl* n, int B, bool grad)
{
  hipStream_t s = n->s;
  if (upload_film(n, B, n->d_film) ||
      forward(n, n->xb, n->ub, nullptr, B, n->d_film, n->out, n->scale, grad ? &n->pre : nullptr,
              s))
    return -1;
  const size_t L = n->layers.size();
  NN_CHECK(unet::loss(n->out[L - 1], n->tb, grad ? n->dy[L - 1] : nullptr, n->lossb, n->Z0,
                      n->cpad, n->zn, n->cout, 1.f / n->out_scale, B, s));
  if (!grad) return 0;
  NN_CHECK(hipMemsetAsync(n->g, 0, n->nparam * 4, s));
  for (size_t k = 0; k + 1 < L; ++k)
    NN_CHECK(hipMemsetAsync(n->dy[k], 0, act_size(n->layers[k], B) * 4, s));
  size_t soff = n->scale_per_sample;
  for (size_t kk = L; kk-- > 0;) {
    const Layer& l = n->layers[kk];
    soff -= l.cout;
    const size_t na = act_size(l, B);
    if (l.skip >= 0)
      hipLaunchKernelGGL(k_add, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s,
                         n->dy[l.skip], n->dy[kk], na);
    NN_CHECK(unet::film_bwd(n->dy[kk], n->pre[kk], n->scale + soff * B, n->ub, n->dpre,
                            n->g + l.a_off, B, l.zout, l.cout, s));
    const float* in = kk == 0 ? n->xb : n->out[kk - 1];
    unet::WgradArgs wa{n->dpre, in, n->g + l.w_off, B, l.zin, l.zout, l.cin, l.cout, l.taps,
                       l.dil, l.mode};
    NN_CHECK(unet::wgrad(wa, s));
    if (kk == 0) continue;
    // d(input) = conv^T(dpre), added into the previous layer's output gradient
    unet::ConvArgs a{};
    a.in = n->dpre;
    a.w = n->wb + l.w_off;
    a.out = n->dy[kk - 1];
    a.skip = n->dy[kk - 1];
    a.B = B; a.Zin = l.zout; a.Zout = l.zin; a.Cin = l.cout; a.Cout = l.cin;
    a.taps = l.taps; a.dil = l.dil;
    a.mode = l.mode == unet::kSame ? unet::kSameFlip
                                   : (l.mode == unet::kDown ? unet::kUp
                                                            : (l.mode == unet::kUp ? unet::kDown
                                                                                   : unet::kPoint));
    NN_CHECK(unet::conv(a, s));
  }
  return 0;
}

// x (host or device) [B][zn_in][cin] rows into xb [B][Z0][cin] (zero padding)
int stage_inputs(NetImpl* n, const float* x, const float* lnk, const float* y, int B,
                 hipMemcpyKind kind)
{
  const size_t row = (size_t)n->nz * n->cin;
  NN_CHECK(hipMemsetAsync(n->xb, 0, (size_t)B * n->Z0 * n->cin * 4, n->s));
  NN_CHECK(hipMemcpy2DAsync(n->xb, (size_t)n->Z0 * n->cin * 4, x, row * 4, row * 4, B, kind, n->s));
  if (y)
    NN_CHECK(hipMemcpyAsync(n->tb, y, (size_t)B * n->zn * n->cout * 4, kind, n->s));
  NN_CHECK(hipMemcpyAsync(n->lossb, lnk, (size_t)B * 4, kind, n->s));
  hipLaunchKernelGGL(k_norm_u, dim3((B + 255) / 256), dim3(256), 0, n->s, n->lossb, nullptr, n->ub,
                     B, n->u_mu, 1.f / n->u_sigma);
  NN_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

NetView net_view(const NetImpl* n)
{
  return NetView{n->nr, n->nz, n->Z0, n->cin, n->cout, n->cpad, n->zn, n->out_scale, n->u_mu,
                 n->u_sigma};
}

int net_infer(NetImpl* n, const float* x, const float* u, const uint32_t* dyn_b, int max_b,
              float** y, hipStream_t s)
{
  if (max_b > n->inf_b) return fail_msg("surrogate inference buffers are not reserved");
  if (forward(n, x, u, dyn_b, n->inf_b, n->d_film_inf, n->iout, n->iscale, nullptr, s))
    return -1;
  *y = n->iout.back();
  return 0;
}

int net_reserve_infer(NetImpl* n, int B)
{
  if (B <= n->inf_b) return 0;
  NN_CHECK(hipStreamSynchronize(n->s));
  free_bufs(n->iout);
  (void)hipFree(n->iscale);
  n->iscale = nullptr;
  n->inf_b = B;
  for (const Layer& l : n->layers) {
    float* p = nullptr;
    NN_CHECK(hipMalloc(&p, act_size(l, B) * 4));
    n->iout.push_back(p);
  }
  NN_CHECK(hipMalloc(&n->iscale, n->scale_per_sample * B * 4));
  return upload_film(n, B, n->d_film_inf);
}

int net_data(NetImpl* n, NetData* out)
{
  *out = n->data;
  return 0;
}

}  // namespace hps

using namespace hps;

struct hps_nn {
  NetImpl impl;
};

namespace hps {
NetImpl* net_of(hps_nn* h) { return h ? &h->impl : nullptr; }
}  // namespace hps

extern "C" {

int hps_nn_create(hps_nn** out, uint32_t nr, uint32_t nz, const uint32_t* widths, uint32_t nbot,
                  uint32_t max_batch, uint32_t seed)
{
  if (!out || !widths) return fail_msg("null argument");
  if (nr == 0 || nr % 16 != 0 || nz == 0 || nz % 64 != 0)
    return fail_msg("the surrogate needs nr a multiple of 16 and nz a multiple of 64");
  for (int k = 0; k < 4; ++k)
    if (widths[k] == 0 || widths[k] % 16 != 0 || widths[k] > (uint32_t)unet::kMaxCout)
      return fail_msg("layer widths must be multiples of 16 of at most 144");
  if (2 * (nr + 1) > (uint32_t)unet::kMaxCout) return fail_msg("nr too large for the output layer");
  if (nbot == 0 || nbot > 6) return fail_msg("1 to 6 bottom layers");
  if (max_batch == 0) return fail_msg("max_batch must be positive");
  hps_nn* h = new hps_nn();
  NetImpl* n = &h->impl;
  n->nr = (int)nr; n->nz = (int)nz; n->Z0 = (int)nz + kPad; n->cin = (int)nr;
  n->cout = 2 * ((int)nr + 1); n->cpad = (n->cout + 15) / 16 * 16; n->zn = (int)nz + 1;
  build_layers(n, widths, nbot);
  for (const Layer& l : n->layers) n->scale_per_sample += l.cout;
  bool ok = hipStreamCreate(&n->s) == hipSuccess;
  const size_t np = n->nparam * 4;
  for (float** p : {&n->w, &n->g, &n->m, &n->v}) ok = ok && hipMalloc(p, np) == hipSuccess;
  ok = ok && hipMalloc(&n->wt, n->nparam * 2) == hipSuccess &&
       hipMalloc(&n->wb, n->nparam * 2) == hipSuccess &&
       hipMalloc(&n->d_film, n->layers.size() * sizeof(unet::FilmLayer)) == hipSuccess &&
       hipMalloc(&n->d_film_inf, n->layers.size() * sizeof(unet::FilmLayer)) == hipSuccess &&
       hipMalloc(&n->data.count, 4) == hipSuccess && hipMemset(n->data.count, 0, 4) == hipSuccess;
  if (!ok) { free_net(n); delete h; return fail_msg("surrogate allocation failed"); }
  // He-style initialisation scaled for a deep linear network: each layer
  // keeps the signal variance (std = 1/sqrt(taps * Cin)); FiLM coefficients 0
  std::vector<float> p(n->nparam, 0.f);
  std::mt19937 rng(seed);
  for (const Layer& l : n->layers) {
    std::normal_distribution<float> nd(0.f, 1.f / std::sqrt((float)(l.taps * l.cin)));
    for (size_t i = 0; i < (size_t)l.cout * l.taps * l.cin; ++i) p[l.w_off + i] = nd(rng);
  }
  if (hipMemcpy(n->w, p.data(), np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, np) != hipSuccess || hipMemset(n->v, 0, np) != hipSuccess ||
      refresh_bf16(n) || alloc_train(n, (int)max_batch)) {
    free_net(n);
    delete h;
    return fail_msg("surrogate initialisation failed");
  }
  *out = h;
  return 0;
}

int hps_nn_destroy(hps_nn* h)
{
  if (!h) return 0;
  (void)hipStreamSynchronize(h->impl.s);
  free_net(&h->impl);
  (void)hipStreamDestroy(h->impl.s);
  delete h;
  return 0;
}

uint32_t hps_nn_num_params(const hps_nn* h) { return h ? (uint32_t)h->impl.nparam : 0; }

int hps_nn_get_params(hps_nn* h, float* host)
{
  if (!h || !host) return fail_msg("null argument");
  NN_CHECK(hipStreamSynchronize(h->impl.s));
  NN_CHECK(hipMemcpy(host, h->impl.w, h->impl.nparam * 4, hipMemcpyDeviceToHost));
  return 0;
}

int hps_nn_set_params(hps_nn* h, const float* host)
{
  if (!h || !host) return fail_msg("null argument");
  NN_CHECK(hipStreamSynchronize(h->impl.s));
  NN_CHECK(hipMemcpy(h->impl.w, host, h->impl.nparam * 4, hipMemcpyHostToDevice));
  if (refresh_bf16(&h->impl)) return -1;
  NN_CHECK(hipStreamSynchronize(h->impl.s));
  return 0;
}

int hps_nn_layers(const hps_nn* h, int32_t* table, uint32_t max_layers, uint32_t* nlayers)
{
  if (!h || !nlayers) return fail_msg("null argument");
  const auto& L = h->impl.layers;
  *nlayers = (uint32_t)L.size();
  for (size_t k = 0; table && k < L.size() && k < max_layers; ++k) {
    const int32_t row[8] = {L[k].mode, L[k].cin, L[k].cout, L[k].taps, L[k].dil, L[k].skip,
                            (int32_t)L[k].w_off, (int32_t)L[k].a_off};
    std::memcpy(table + 8 * k, row, sizeof row);
  }
  return 0;
}

int hps_nn_dims(const hps_nn* h, uint32_t* out)
{
  if (!h || !out) return fail_msg("null argument");
  const NetImpl& n = h->impl;
  const uint32_t d[7] = {(uint32_t)n.Z0, (uint32_t)n.cin, (uint32_t)n.cout, (uint32_t)n.cpad,
                         (uint32_t)n.zn, (uint32_t)n.nr, (uint32_t)n.nz};
  std::memcpy(out, d, sizeof d);
  return 0;
}

int hps_nn_set_norm(hps_nn* h, float out_scale, float u_mu, float u_sigma)
{
  if (!h) return fail_msg("null argument");
  if (!(out_scale > 0.f) || !(u_sigma > 0.f)) return fail_msg("scales must be positive");
  h->impl.out_scale = out_scale; h->impl.u_mu = u_mu; h->impl.u_sigma = u_sigma;
  return 0;
}

int hps_nn_get_norm(const hps_nn* h, float* out)
{
  if (!h || !out) return fail_msg("null argument");
  out[0] = h->impl.out_scale; out[1] = h->impl.u_mu; out[2] = h->impl.u_sigma;
  return 0;
}

int hps_nn_data_reserve(hps_nn* h, uint32_t cap)
{
  if (!h) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  NN_CHECK(hipStreamSynchronize(n->s));
  uint32_t have = 0;
  NN_CHECK(hipMemcpy(&have, n->data.count, 4, hipMemcpyDeviceToHost));
  have = std::min(have, n->data.cap);
  if (cap < have) return fail_msg("the dataset holds more samples than the new capacity");
  float *x = nullptr, *y = nullptr, *u = nullptr;
  const size_t rx = (size_t)n->nz * n->cin, ry = (size_t)n->zn * n->cout;
  NN_CHECK(hipMalloc(&x, (size_t)cap * rx * 4));
  NN_CHECK(hipMalloc(&y, (size_t)cap * ry * 4));
  NN_CHECK(hipMalloc(&u, (size_t)cap * 4));
  if (have) {
    NN_CHECK(hipMemcpy(x, n->data.x, have * rx * 4, hipMemcpyDeviceToDevice));
    NN_CHECK(hipMemcpy(y, n->data.y, have * ry * 4, hipMemcpyDeviceToDevice));
    NN_CHECK(hipMemcpy(u, n->data.u, have * 4, hipMemcpyDeviceToDevice));
  }
  (void)hipFree(n->data.x); (void)hipFree(n->data.y); (void)hipFree(n->data.u);
  n->data.x = x; n->data.y = y; n->data.u = u; n->data.cap = cap;
  NN_CHECK(hipMemcpy(n->data.count, &have, 4, hipMemcpyHostToDevice));
  return 0;
}

int hps_nn_data_size(hps_nn* h, uint32_t* out)
{
  if (!h || !out) return fail_msg("null argument");
  NN_CHECK(hipDeviceSynchronize());
  uint32_t c = 0;
  NN_CHECK(hipMemcpy(&c, h->impl.data.count, 4, hipMemcpyDeviceToHost));
  *out = std::min(c, h->impl.data.cap);
  return 0;
}

int hps_nn_data_append(hps_nn* h, const float* x, const float* y, const float* lnk, uint32_t cnt)
{
  if (!h || !x || !y || !lnk) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  uint32_t have = 0;
  if (hps_nn_data_size(h, &have)) return -1;
  if (have + cnt > n->data.cap) return fail_msg("the dataset is full (hps_nn_data_reserve)");
  const size_t rx = (size_t)n->nz * n->cin, ry = (size_t)n->zn * n->cout;
  NN_CHECK(hipMemcpy(n->data.x + have * rx, x, cnt * rx * 4, hipMemcpyHostToDevice));
  NN_CHECK(hipMemcpy(n->data.y + have * ry, y, cnt * ry * 4, hipMemcpyHostToDevice));
  NN_CHECK(hipMemcpy(n->data.u + have, lnk, cnt * 4, hipMemcpyHostToDevice));
  have += cnt;
  NN_CHECK(hipMemcpy(n->data.count, &have, 4, hipMemcpyHostToDevice));
  return 0;
}

int hps_nn_data_shuffle(hps_nn* h, uint32_t seed)
{
  if (!h) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  uint32_t have = 0;
  if (hps_nn_data_size(h, &have)) return -1;
  if (have < 2) return 0;
  std::vector<uint32_t> perm(have);
  for (uint32_t i = 0; i < have; ++i) perm[i] = i;
  std::mt19937 rng(seed);
  std::shuffle(perm.begin(), perm.end(), rng);
  const size_t rx = (size_t)n->nz * n->cin, ry = (size_t)n->zn * n->cout;
  uint32_t* d_perm = nullptr;
  float *x = nullptr, *y = nullptr, *u = nullptr;
  NN_CHECK(hipMalloc(&d_perm, have * 4));
  NN_CHECK(hipMalloc(&x, (size_t)n->data.cap * rx * 4));
  NN_CHECK(hipMalloc(&y, (size_t)n->data.cap * ry * 4));
  NN_CHECK(hipMalloc(&u, (size_t)n->data.cap * 4));
  NN_CHECK(hipMemcpy(d_perm, perm.data(), have * 4, hipMemcpyHostToDevice));
  NN_CHECK(unet::gather(n->data.x, d_perm, x, (int)have, (int)rx, n->s));
  NN_CHECK(unet::gather(n->data.y, d_perm, y, (int)have, (int)ry, n->s));
  NN_CHECK(unet::gather(n->data.u, d_perm, u, (int)have, 1, n->s));
  NN_CHECK(hipStreamSynchronize(n->s));
  (void)hipFree(n->data.x); (void)hipFree(n->data.y); (void)hipFree(n->data.u);
  (void)hipFree(d_perm);
  n->data.x = x; n->data.y = y; n->data.u = u;
  return 0;
}

int hps_nn_data_get(hps_nn* h, uint32_t first, uint32_t cnt, float* x, float* y, float* lnk)
{
  if (!h) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  uint32_t have = 0;
  if (hps_nn_data_size(h, &have)) return -1;
  if ((uint64_t)first + cnt > have) return fail_msg("dataset range out of bounds");
  const size_t rx = (size_t)n->nz * n->cin, ry = (size_t)n->zn * n->cout;
  if (x) NN_CHECK(hipMemcpy(x, n->data.x + first * rx, cnt * rx * 4, hipMemcpyDeviceToHost));
  if (y) NN_CHECK(hipMemcpy(y, n->data.y + first * ry, cnt * ry * 4, hipMemcpyDeviceToHost));
  if (lnk) NN_CHECK(hipMemcpy(lnk, n->data.u + first, cnt * 4, hipMemcpyDeviceToHost));
  return 0;
}

int hps_nn_loss_grad(hps_nn* h, const float* x, const float* lnk, const float* y, uint32_t count,
                     float* loss, float* grad)
{
  if (!h || !x || !lnk || !y || !loss) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  if ((int)count > n->max_b) return fail_msg("more samples than max_batch");
  if (stage_inputs(n, x, lnk, y, (int)count, hipMemcpyHostToDevice) ||
      loss_grad(n, (int)count, grad != nullptr))
    return -1;
  NN_CHECK(hipStreamSynchronize(n->s));
  NN_CHECK(hipMemcpy(loss, n->lossb, count * 4, hipMemcpyDeviceToHost));
  if (grad) NN_CHECK(hipMemcpy(grad, n->g, n->nparam * 4, hipMemcpyDeviceToHost));
  return 0;
}

int hps_nn_predict(hps_nn* h, const float* x, const float* lnk, uint32_t count, float* y)
{
  if (!h || !x || !lnk || !y) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  for (uint32_t c0 = 0; c0 < count; c0 += (uint32_t)n->max_b) {
    const int B = (int)std::min<uint32_t>((uint32_t)n->max_b, count - c0);
    if (stage_inputs(n, x + (size_t)c0 * n->nz * n->cin, lnk + c0, nullptr, B,
                     hipMemcpyHostToDevice) ||
        upload_film(n, B, n->d_film) ||
        forward(n, n->xb, n->ub, nullptr, B, n->d_film, n->out, n->scale, nullptr, n->s))
      return -1;
    // rows of the node positions, channels of the field, times the output scale
    const float* o = n->out.back();
    std::vector<float> tmp((size_t)B * n->Z0 * n->cpad);
    NN_CHECK(hipStreamSynchronize(n->s));
    NN_CHECK(hipMemcpy(tmp.data(), o, tmp.size() * 4, hipMemcpyDeviceToHost));
    for (int b = 0; b < B; ++b)
      for (int z = 0; z < n->zn; ++z)
        for (int c = 0; c < n->cout; ++c)
          y[(((size_t)c0 + b) * n->zn + z) * n->cout + c] =
              tmp[((size_t)b * n->Z0 + z) * n->cpad + c] * n->out_scale;
  }
  return 0;
}

int hps_nn_eval(hps_nn* h, uint32_t first, uint32_t count, float* loss)
{
  if (!h || !loss) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  uint32_t have = 0;
  if (hps_nn_data_size(h, &have)) return -1;
  if ((uint64_t)first + count > have) return fail_msg("dataset range out of bounds");
  const size_t rx = (size_t)n->nz * n->cin, ry = (size_t)n->zn * n->cout;
  for (uint32_t c0 = 0; c0 < count; c0 += (uint32_t)n->max_b) {
    const int B = (int)std::min<uint32_t>((uint32_t)n->max_b, count - c0);
    const uint32_t s0 = first + c0;
    if (stage_inputs(n, n->data.x + s0 * rx, n->data.u + s0, n->data.y + s0 * ry, B,
                     hipMemcpyDeviceToDevice) ||
        loss_grad(n, B, false))
      return -1;
    NN_CHECK(hipStreamSynchronize(n->s));
    NN_CHECK(hipMemcpy(loss + c0, n->lossb, B * 4, hipMemcpyDeviceToHost));
  }
  return 0;
}

int hps_nn_train(hps_nn* h, uint32_t first, uint32_t count, uint32_t steps, uint32_t batch,
                 float lr, float lr_end, uint32_t seed, float* losses)
{
  if (!h) return fail_msg("null argument");
  NetImpl* n = &h->impl;
  if ((int)batch > n->max_b || batch == 0) return fail_msg("batch must be in 1..max_batch");
  uint32_t have = 0;
  if (hps_nn_data_size(h, &have)) return -1;
  if (count == 0 || (uint64_t)first + count > have) return fail_msg("dataset range out of bounds");
  if ((int)steps > n->hist_cap) {
    (void)hipFree(n->hist);
    NN_CHECK(hipMalloc(&n->hist, (size_t)steps * 4));
    n->hist_cap = (int)steps;
  }
  if (!n->idx) NN_CHECK(hipMalloc(&n->idx, (size_t)n->max_b * 4));
  const size_t rx = (size_t)n->nz * n->cin, ry = (size_t)n->zn * n->cout;
  std::mt19937 rng(seed);
  std::uniform_int_distribution<uint32_t> pick(first, first + count - 1);
  std::vector<uint32_t> idx(batch);
  const int B = (int)batch;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t st = 0; st < steps; ++st) {
    for (auto& i : idx) i = pick(rng);
    NN_CHECK(hipMemcpyAsync(n->idx, idx.data(), batch * 4, hipMemcpyHostToDevice, n->s));
    // inputs: the nz bin rows into the zero-padded Z0 rows
    NN_CHECK(hipMemsetAsync(n->xb, 0, (size_t)B * n->Z0 * n->cin * 4, n->s));
    NN_CHECK(unet::gather(n->data.x, n->idx, n->tb, B, (int)rx, n->s));   // staging in tb
    NN_CHECK(hipMemcpy2DAsync(n->xb, (size_t)n->Z0 * n->cin * 4, n->tb, rx * 4, rx * 4, B,
                              hipMemcpyDeviceToDevice, n->s));
    NN_CHECK(unet::gather(n->data.y, n->idx, n->tb, B, (int)ry, n->s));
    hipLaunchKernelGGL(k_norm_u, dim3((B + 255) / 256), dim3(256), 0, n->s, n->data.u, n->idx,
                       n->ub, B, n->u_mu, 1.f / n->u_sigma);
    if (loss_grad(n, B, true)) return -1;
    ++n->step;
    // cosine decay from lr to lr_end over the call
    const float frac = steps > 1 ? (float)st / (float)(steps - 1) : 0.f;
    const float lr_now = lr_end + 0.5f * (lr - lr_end) * (1.f + std::cos(3.14159265f * frac));
    const float lr_t = lr_now * std::sqrt(1.f - std::pow(b2, (float)n->step)) /
                       (1.f - std::pow(b1, (float)n->step));
    NN_CHECK(unet::adam(n->w, n->g, n->m, n->v, (int)n->nparam, lr_t, b1, b2, eps, n->s));
    if (refresh_bf16(n)) return -1;
    hipLaunchKernelGGL(k_mean, dim3(1), dim3(64), 0, n->s, n->lossb, B, n->hist + st);
  }
  NN_CHECK(hipStreamSynchronize(n->s));
  if (losses) NN_CHECK(hipMemcpy(losses, n->hist, steps * 4, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
