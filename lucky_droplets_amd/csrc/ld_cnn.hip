// C-ABI of the learned subgrid form (declared in include/ld.h): a periodic
// 3-D CNN of nlayers 3x3x3 convolutions (ReLU between them), bf16 weights and
// activations, FP32 accumulation and FP32 master parameters trained with Adam.
#include "ld_cnn.hpp"
#include "ld_engine_base.hpp"
#include <new>
#include <vector>

using namespace ld::cnn;

// blocks of the loss kernel (one FP64 partial sum each)
#define LD_CNN_MSE_BLOCKS 2048
using ld::set_error;

struct ld_cnn {
  int N = 0, L = 0, cin = 0, cout = 0;
  int ch[LD_CNN_MAX_LAYERS + 1] = {};     // padded channels at each layer boundary
  size_t nodes = 0;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  int64_t nparams = 0;
  int64_t off[LD_CNN_MAX_LAYERS] = {};      // layer l: W[27][co][ci] then b[co]
  bf16* Wb[LD_CNN_MAX_LAYERS] = {};
  bf16* Wt[LD_CNN_MAX_LAYERS] = {};
  bf16* act[LD_CNN_MAX_LAYERS] = {};        // act[l]: input of layer l
  bf16* dz[2] = {nullptr, nullptr};
  float* yp = nullptr;                      // FP32 output [node][ch[L]]
  float* part = nullptr;
  double* mpart = nullptr;
  int nblk_w = 0, chunk = 0;
  int nblk_w2 = 0, chunk2 = 0;              // k_wgrad2 blocks and nodes per block (N >= 16)
  bf16 *xt = nullptr, *dzt = nullptr;       // node-blocked channel-major copies for the weight gradient
  bf16* actB[LD_CNN_MAX_LAYERS] = {};       // node-blocked copies of the hidden activations (N >= 16)
  int adam_t = 0;
  bool packed = false;
  ~ld_cnn() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* a : allocs) (void)hipFree(a);
  }
  template <class X> int alloc(X*& p, size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, (n ? n : 1) * sizeof(X)) != hipSuccess) { set_error("hipMalloc failed in ld_cnn"); return LD_ERR_HIP; }
    allocs.push_back(q);
    p = (X*)q;
    return LD_OK;
  }
};

#define CK(x)                       \
  do {                              \
    int rc_ = (x);                  \
    if (rc_ != LD_OK) return rc_;   \
  } while (0)
#define HK(call)                                                              \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error("%s failed: %s", #call, hipGetErrorString(e_));               \
      return LD_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)
#define LAUNCHED() HK(hipGetLastError())

namespace {

// Convolution launch for the channel pairs the network uses (16 or 32 each).
// gfx950 double-rate form: NT node tiles of 16 per wavefront (a z-line
// segment of 16 NT nodes), persistent blocks of WPB wavefronts.
template <int CI, int CO, bool FLIP, bool RELU, bool MASK>
int conv2(ld_cnn* c, const bf16* X, const bf16* W, const float* b, const bf16* mask, bf16* Yb, float* Yf,
          bf16* Yblk) {
  const int N = c->N;
  // 4 node tiles per wavefront and 8 wavefronts per block (sharing one LDS
  // copy of the weights): measured at 128^3, channels 3-32-32-32-3, forward
  // 0.76 -> 0.67 ms against 8 tiles and 4 wavefronts (2 tiles: 0.75 ms;
  // 8 tiles with 8 wavefronts: 0.74 ms)
  constexpr int WPB = 8;
  const int NT = N >= 64 ? 4 : N / 16;
  const int nseg = (int)(c->nodes / (16 * NT));
  int G = (nseg + WPB - 1) / WPB;
  if (G > 2048) G = 2048;
  const dim3 g((unsigned)G), t(64 * WPB);
#define LD_NT_CASE(K)                                                                                            \
  if (NT == K) {                                                                                                 \
    if constexpr (CI == 32)                                                                                      \
      hipLaunchKernelGGL((k_conv3<CO, FLIP, RELU, MASK, K, WPB>), g, t, 0, c->st, X, W, b, mask, Yb, Yf, N, nseg,  \
                         Yblk);                                                                                  \
    else                                                                                                         \
      hipLaunchKernelGGL((k_conv2<CI, CO, FLIP, RELU, MASK, K, WPB>), g, t, 0, c->st, X, W, b, mask, Yb, Yf, N,    \
                         nseg, Yblk);                                                                            \
    LAUNCHED();                                                                                                  \
    return LD_OK;                                                                                                \
  }
  LD_NT_CASE(1) LD_NT_CASE(2) LD_NT_CASE(4)
#undef LD_NT_CASE
  set_error("conv2: unsupported N %d", N);
  return LD_ERR_ARG;
}

// Yblk (N >= 16 only; may be null): a second copy of Yb in the node-blocked
// layout the weight gradient reads.
template <bool FLIP, bool RELU, bool MASK>
int conv(ld_cnn* c, int CI, int CO, const bf16* X, const bf16* W, const float* b, const bf16* mask, bf16* Yb,
         float* Yf, bf16* Yblk = nullptr) {
  if (c->N >= 16) {
    if (CI == 16 && CO == 16) return conv2<16, 16, FLIP, RELU, MASK>(c, X, W, b, mask, Yb, Yf, Yblk);
    if (CI == 16 && CO == 32) return conv2<16, 32, FLIP, RELU, MASK>(c, X, W, b, mask, Yb, Yf, Yblk);
    if (CI == 32 && CO == 16) return conv2<32, 16, FLIP, RELU, MASK>(c, X, W, b, mask, Yb, Yf, Yblk);
    if (CI == 32 && CO == 32) return conv2<32, 32, FLIP, RELU, MASK>(c, X, W, b, mask, Yb, Yf, Yblk);
  }
  const dim3 g((unsigned)(c->nodes / 64)), t(256);
#define LD_CONV_CASE(A, B)                                                                                    \
  if (CI == A && CO == B) {                                                                                   \
    hipLaunchKernelGGL((k_conv<A, B, FLIP, RELU, MASK>), g, t, 0, c->st, X, W, b, mask, Yb, Yf, c->N);      \
    LAUNCHED();                                                                                               \
    return LD_OK;                                                                                             \
  }
  LD_CONV_CASE(16, 16) LD_CONV_CASE(16, 32) LD_CONV_CASE(32, 16) LD_CONV_CASE(32, 32)
#undef LD_CONV_CASE
  set_error("unsupported channel pair %d -> %d", CI, CO);
  return LD_ERR_ARG;
}

template <int C> int cmajor(ld_cnn* c, const bf16* A, bf16* AT) {
  const size_t n = c->nodes / 8 * (C / 8);
  hipLaunchKernelGGL((k_cblock<C>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, A, AT, c->nodes);
  LAUNCHED();
  return LD_OK;
}

// Weight and bias gradients of one layer into c->part (per-block partials);
// returns the number of partial blocks in *nblk.  XB / dZB: node-blocked
// copies of X / dZ when a producing kernel already wrote them (N >= 16), else
// null and they are made here.
int wgrad(ld_cnn* c, int CI, int CO, const bf16* X, const bf16* dZ, const bf16* XB, const bf16* dZB, int* nblk) {
  if (c->N >= 16) {
    if (!XB) { CK(CI == 16 ? cmajor<16>(c, X, c->xt) : cmajor<32>(c, X, c->xt)); XB = c->xt; }
    if (!dZB) { CK(CO == 16 ? cmajor<16>(c, dZ, c->dzt) : cmajor<32>(c, dZ, c->dzt)); dZB = c->dzt; }
    // one neighbour line per wavefront (three measured slower: 396 -> 411 us at 128^3, 32 -> 32)
    constexpr int LPW = 1;
    const dim3 g((unsigned)c->nblk_w2), t(64 * 9 / LPW);
#define LD_WG2_CASE(A, B)                                                                                  \
    if (CI == A && CO == B) {                                                                              \
      hipLaunchKernelGGL((k_wgrad2<A, B, LPW>), g, t, 0, c->st, XB, dZB, c->part, c->N, c->chunk2);          \
      LAUNCHED();                                                                                          \
      *nblk = c->nblk_w2;                                                                                  \
      return LD_OK;                                                                                        \
    }
    LD_WG2_CASE(16, 16) LD_WG2_CASE(16, 32) LD_WG2_CASE(32, 16) LD_WG2_CASE(32, 32)
#undef LD_WG2_CASE
    set_error("unsupported channel pair %d -> %d", CI, CO);
    return LD_ERR_ARG;
  }
  *nblk = c->nblk_w;
  const dim3 g((unsigned)c->nblk_w), t(256);
#define LD_WG_CASE(A, B)                                                                             \
  if (CI == A && CO == B) {                                                                          \
    hipLaunchKernelGGL((k_wgrad<A, B>), g, t, 0, c->st, X, dZ, c->part, c->N, c->chunk);            \
    LAUNCHED();                                                                                      \
    return LD_OK;                                                                                    \
  }
  LD_WG_CASE(16, 16) LD_WG_CASE(16, 32) LD_WG_CASE(32, 16) LD_WG_CASE(32, 32)
#undef LD_WG_CASE
  set_error("unsupported channel pair %d -> %d", CI, CO);
  return LD_ERR_ARG;
}

int pack_weights(ld_cnn* c) {
  for (int l = 0; l < c->L; ++l) {
    hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, c->st, (const float*)(c->params + c->off[l]),
                       c->ch[l], c->ch[l + 1], c->Wb[l], c->Wt[l]);
    LAUNCHED();
  }
  c->packed = true;
  return LD_OK;
}

// Forward pass from the FP32 input rows x[node][cin]; keeps every layer's
// input for the backward pass; FP32 output in c->yp.  for_grad: the hidden
// layers also write node-blocked copies of their outputs (c->actB, N >= 16)
// for the weight gradient.
int forward(ld_cnn* c, const float* x, bool for_grad) {
  if (!c->packed) CK(pack_weights(c));
  const size_t gp = (c->nodes + 255) / 256;
  hipLaunchKernelGGL(k_pack_input, dim3((unsigned)(gp < (1u << 20) ? gp : (1u << 20))), dim3(256), 0, c->st, x, c->cin,
                     c->ch[0], c->nodes, c->act[0]);
  LAUNCHED();
  for (int l = 0; l < c->L; ++l) {
    const float* b = c->params + c->off[l] + 27LL * c->ch[l + 1] * c->ch[l];
    if (l + 1 < c->L)
      CK((conv<false, true, false>(c, c->ch[l], c->ch[l + 1], c->act[l], c->Wb[l], b, nullptr, c->act[l + 1], nullptr,
                                   for_grad ? c->actB[l + 1] : nullptr)));
    else
      CK((conv<false, false, false>(c, c->ch[l], c->ch[l + 1], c->act[l], c->Wb[l], b, nullptr, nullptr, c->yp)));
  }
  return LD_OK;
}

// Mean-squared-error loss against target[node][cout] and the parameter
// gradients (in c->grads) by backpropagation through the stored activations.
int backward(ld_cnn* c, const float* target, double* loss) {
  const int nb = LD_CNN_MSE_BLOCKS;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, c->st, (const float*)c->yp, target, c->cout, c->ch[c->L], c->nodes,
                     c->dz[0], c->mpart);
  LAUNCHED();
  int cur = 0;
  bool dzb = false;        // c->dzt holds the node-blocked copy of dz[cur]
  for (int l = c->L - 1; l >= 0; --l) {
    const int CI = c->ch[l], CO = c->ch[l + 1];
    int nblk = 0;
    CK(wgrad(c, CI, CO, c->act[l], c->dz[cur], l > 0 ? c->actB[l] : nullptr, dzb ? c->dzt : nullptr, &nblk));
    const int n = 27 * CO * CI + CO;
    hipLaunchKernelGGL(k_reduce_parts, dim3((n + 255) / 256), dim3(256), 0, c->st, (const float*)c->part, nblk, n,
                       c->grads + c->off[l]);
    LAUNCHED();
    if (l > 0) {
      // data gradient through the convolution (transposed weights, mirrored taps)
      // and the ReLU of layer l - 1 (whose output is act[l])
      CK((conv<true, false, true>(c, CO, CI, c->dz[cur], c->Wt[l], nullptr, c->act[l], c->dz[cur ^ 1], nullptr,
                                  c->N >= 16 ? c->dzt : nullptr)));
      dzb = c->N >= 16;
      cur ^= 1;
    }
  }
  if (loss) {
    std::vector<double> h(nb);
    HK(hipMemcpyAsync(h.data(), c->mpart, nb * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HK(hipStreamSynchronize(c->st));
    double s = 0.0;
    for (double v : h) s += v;
    *loss = s / ((double)c->nodes * c->cout);
  }
  return LD_OK;
}

}  // namespace

extern "C" {

int ld_cnn_create(const ld_cnn_params* p, ld_cnn** out) {
  if (!p || !out) { set_error("null argument"); return LD_ERR_ARG; }
  *out = nullptr;
  if (p->N < 4 || p->N > 512 || (p->N & (p->N - 1))) { set_error("N must be a power of two in [4, 512]"); return LD_ERR_ARG; }
  if (p->nlayers < 1 || p->nlayers > LD_CNN_MAX_LAYERS) { set_error("nlayers must be in [1, %d]", LD_CNN_MAX_LAYERS); return LD_ERR_ARG; }
  ld_cnn* c = new (std::nothrow) ld_cnn;
  if (!c) { set_error("out of host memory"); return LD_ERR_ARG; }
  c->N = p->N;
  c->L = p->nlayers;
  c->nodes = (size_t)p->N * p->N * p->N;
  c->cin = p->channels[0];
  c->cout = p->channels[p->nlayers];
  for (int l = 0; l <= c->L; ++l) {
    const int k = p->channels[l];
    if (k < 1 || k > 32 || ((l > 0 && l < c->L) && k != 16 && k != 32)) {
      set_error("channels[%d] = %d: inputs/outputs 1..32, hidden layers 16 or 32", l, k);
      delete c;
      return LD_ERR_ARG;
    }
    c->ch[l] = k <= 16 ? 16 : 32;
  }
  int rc = LD_OK;
  int64_t np = 0;
  for (int l = 0; l < c->L; ++l) {
    c->off[l] = np;
    np += 27LL * c->ch[l] * c->ch[l + 1] + c->ch[l + 1];
  }
  c->nparams = np;
  // weight-gradient blocks: at most 256, each over a multiple of 64 nodes
  c->chunk = 64;
  while (c->nodes / c->chunk > 256) c->chunk *= 2;
  c->nblk_w = (int)(c->nodes / c->chunk);
  // k_wgrad2: at most 512 blocks, each over a multiple of 32 nodes
  c->chunk2 = 32;
  while (c->nodes / c->chunk2 > 512) c->chunk2 *= 2;
  c->nblk_w2 = (int)(c->nodes / c->chunk2);
  int64_t maxlayer = 0;
  for (int l = 0; l < c->L; ++l) {
    const int64_t n = 27LL * c->ch[l] * c->ch[l + 1] + c->ch[l + 1];
    maxlayer = maxlayer > n ? maxlayer : n;
  }
  if (rc == LD_OK) rc = c->alloc(c->params, np);
  if (rc == LD_OK) rc = c->alloc(c->grads, np);
  if (rc == LD_OK) rc = c->alloc(c->m, np);
  if (rc == LD_OK) rc = c->alloc(c->v, np);
  for (int l = 0; l < c->L && rc == LD_OK; ++l) {
    rc = c->alloc(c->Wb[l], 27 * (size_t)c->ch[l] * c->ch[l + 1]);
    if (rc == LD_OK) rc = c->alloc(c->Wt[l], 27 * (size_t)c->ch[l] * c->ch[l + 1]);
    if (rc == LD_OK) rc = c->alloc(c->act[l], c->nodes * c->ch[l]);
  }
  if (rc == LD_OK) rc = c->alloc(c->dz[0], c->nodes * 32);
  if (rc == LD_OK) rc = c->alloc(c->dz[1], c->nodes * 32);
  if (rc == LD_OK) rc = c->alloc(c->yp, c->nodes * c->ch[c->L]);
  if (rc == LD_OK) rc = c->alloc(c->part, (size_t)(c->nblk_w > c->nblk_w2 ? c->nblk_w : c->nblk_w2) * maxlayer);
  if (rc == LD_OK && c->N >= 16) rc = c->alloc(c->xt, c->nodes * 32);
  if (rc == LD_OK && c->N >= 16) rc = c->alloc(c->dzt, c->nodes * 32);
  for (int l = 1; l < c->L && rc == LD_OK && c->N >= 16; ++l) rc = c->alloc(c->actB[l], c->nodes * c->ch[l]);
  if (rc == LD_OK) rc = c->alloc(c->mpart, LD_CNN_MSE_BLOCKS);
  if (rc == LD_OK && hipMemset(c->params, 0, np * sizeof(float)) != hipSuccess) rc = LD_ERR_HIP;
  if (rc == LD_OK && hipMemset(c->m, 0, np * sizeof(float)) != hipSuccess) rc = LD_ERR_HIP;
  if (rc == LD_OK && hipMemset(c->v, 0, np * sizeof(float)) != hipSuccess) rc = LD_ERR_HIP;
  if (rc != LD_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return LD_OK;
}

int ld_cnn_destroy(ld_cnn* c) {
  delete c;
  return LD_OK;
}

int ld_cnn_set_stream(ld_cnn* c, void* s) {
  if (!c) { set_error("null ld_cnn"); return LD_ERR_ARG; }
  c->st = (hipStream_t)s;
  return LD_OK;
}

int ld_cnn_nparams(ld_cnn* c, int64_t* n) {
  if (!c || !n) { set_error("null argument"); return LD_ERR_ARG; }
  *n = c->nparams;
  return LD_OK;
}

int ld_cnn_set_params(ld_cnn* c, const float* params) {
  if (!c || !params) { set_error("null argument"); return LD_ERR_ARG; }
  HK(hipMemcpyAsync(c->params, params, c->nparams * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  HK(hipMemsetAsync(c->m, 0, c->nparams * sizeof(float), c->st));
  HK(hipMemsetAsync(c->v, 0, c->nparams * sizeof(float), c->st));
  c->adam_t = 0;
  c->packed = false;
  return LD_OK;
}

int ld_cnn_get_params(ld_cnn* c, float* params) {
  if (!c || !params) { set_error("null argument"); return LD_ERR_ARG; }
  HK(hipMemcpyAsync(params, c->params, c->nparams * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  return LD_OK;
}

int ld_cnn_forward(ld_cnn* c, const float* x, float* y) {
  if (!c || !x || !y) { set_error("null argument"); return LD_ERR_ARG; }
  CK(forward(c, x, false));
  hipLaunchKernelGGL(k_unpack_output, dim3(1024), dim3(256), 0, c->st, (const float*)c->yp, c->ch[c->L], c->cout,
                     c->nodes, y);
  LAUNCHED();
  return LD_OK;
}

int ld_cnn_grad(ld_cnn* c, const float* x, const float* target, float* grads, double* loss) {
  if (!c || !x || !target) { set_error("null argument"); return LD_ERR_ARG; }
  CK(forward(c, x, true));
  CK(backward(c, target, loss));
  if (grads) HK(hipMemcpyAsync(grads, c->grads, c->nparams * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  return LD_OK;
}

int ld_cnn_train_step(ld_cnn* c, const float* x, const float* target, float lr, float beta1, float beta2, float eps,
                      double* loss) {
  if (!c || !x || !target) { set_error("null argument"); return LD_ERR_ARG; }
  CK(forward(c, x, true));
  CK(backward(c, target, loss));
  c->adam_t += 1;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((c->nparams + 255) / 256)), dim3(256), 0, c->st, c->params,
                     (const float*)c->grads, c->m, c->v, (int)c->nparams, lr, beta1, beta2, eps, c->adam_t);
  LAUNCHED();
  c->packed = false;
  return LD_OK;
}

}  // extern "C"
