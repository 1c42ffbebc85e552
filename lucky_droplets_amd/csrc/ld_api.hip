// C-ABI entry points of libld.so (declared in include/ld.h).
#include "ld_engine_base.hpp"
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

struct ld_ctx {
  ld::EngineBase* eng = nullptr;
  ld_params p{};
  int device = 0;     // the HIP device current at ld_create
};

// Scratch of ld_spectrum_noalloc: the per-block shell partial sums of the
// spectrum (256 blocks x at most 257 shells, N <= 512: 526,336 bytes).  Each
// GPU holds its own copy for the life of the process; its device address is
// looked up once per device at ld_create and kept in g_scratch[].  Every call
// on one GPU shares that GPU's copy, so such calls must not overlap.
static constexpr size_t LD_SCRATCH_DOUBLES = 256 * 257;
static __device__ __align__(256) double scratch[LD_SCRATCH_DOUBLES];
static constexpr int LD_MAX_DEVICES = 64;
static double* g_scratch[LD_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

namespace ld {
static thread_local char g_err[1024] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace ld

using ld::set_error;

// Device address of the scratch array on the current device, looked up once
// per device (the array's host-side name is not a device address).
static int scratch_setup(int* device_out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { set_error("hipGetDevice failed: %s", hipGetErrorString(e)); return LD_ERR_HIP; }
  if (dev < 0 || dev >= LD_MAX_DEVICES) { set_error("device %d beyond the %d supported", dev, LD_MAX_DEVICES); return LD_ERR_ARG; }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return LD_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  *device_out = dev;
  return LD_OK;
}

#define LD_CHECK_CTX(c)                                 \
  do {                                                  \
    if (!(c) || !(c)->eng) {                            \
      set_error("null or destroyed ld_ctx");            \
      return LD_ERR_ARG;                                \
    }                                                   \
  } while (0)
#define LD_NONNULL(ptr)                                 \
  do {                                                  \
    if (!(ptr)) {                                       \
      set_error("%s: null pointer argument " #ptr, __func__); \
      return LD_ERR_ARG;                                \
    }                                                   \
  } while (0)

// Field arrays (real and spectral) are moved as 16-byte (FP32) or 32-byte
// (FP64) vectors: a non-null field pointer must be 32-byte aligned.
#define LD_ALIGNED(ptr)                                 \
  do {                                                  \
    if ((uintptr_t)(ptr) % 32 != 0) {                   \
      set_error("%s: field array " #ptr " must be 32-byte aligned", __func__); \
      return LD_ERR_ARG;                                \
    }                                                   \
  } while (0)
#define LD_FIELD(ptr) do { LD_NONNULL(ptr); LD_ALIGNED(ptr); } while (0)

// Entry points that work on one realization reject an ensemble context.
#define LD_ONE(c)                                                                        \
  do {                                                                                   \
    if ((c)->eng->batch() != 1) {                                                        \
      set_error("%s works on one realization; this context holds %d", __func__, (c)->eng->batch()); \
      return LD_ERR_ARG;                                                                 \
    }                                                                                    \
  } while (0)

static ld::EngineBase* make_engine(const ld_params& p) {
  using namespace ld;
#define LD_CASE(TN, NN) \
  if (p.N == NN) return make_engine_##TN##_##NN(p);
  if (p.precision == LD_F32) {
    LD_CASE(f32, 16) LD_CASE(f32, 32) LD_CASE(f32, 64) LD_CASE(f32, 128) LD_CASE(f32, 256) LD_CASE(f32, 512)
  } else if (p.precision == LD_F64) {
    LD_CASE(f64, 16) LD_CASE(f64, 32) LD_CASE(f64, 64) LD_CASE(f64, 128) LD_CASE(f64, 256) LD_CASE(f64, 512)
  }
#undef LD_CASE
  return nullptr;
}

extern "C" {

const char* ld_version(void) { return "ld 0.1 gfx950"; }

int ld_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return LD_ERR_ARG;
  strncpy(buf, ld::g_err, len - 1);
  buf[len - 1] = '\0';
  return LD_OK;
}

int ld_create(const ld_params* params, ld_ctx** out) {
  LD_NONNULL(params);
  LD_NONNULL(out);
  *out = nullptr;
  if (params->precision != LD_F32 && params->precision != LD_F64) {
    set_error("precision must be LD_F32 or LD_F64");
    return LD_ERR_ARG;
  }
  ld::EngineBase* e = make_engine(*params);
  if (!e) {
    set_error("unsupported grid size N=%d (power of two in [16, 512])", params->N);
    return LD_ERR_ARG;
  }
  int dev = 0;
  int rc = scratch_setup(&dev);
  if (rc != LD_OK) {
    delete e;
    return rc;
  }
  rc = e->init();
  if (rc != LD_OK) {
    delete e;
    return rc;
  }
  ld_ctx* c = new (std::nothrow) ld_ctx;
  if (!c) {
    delete e;
    set_error("out of host memory");
    return LD_ERR_ARG;
  }
  c->eng = e;
  c->p = *params;
  c->device = dev;
  *out = c;
  return LD_OK;
}

int ld_destroy(ld_ctx* ctx) {
  if (!ctx) return LD_OK;
  delete ctx->eng;
  ctx->eng = nullptr;
  delete ctx;
  return LD_OK;
}

int ld_set_stream(ld_ctx* ctx, void* s) {
  LD_CHECK_CTX(ctx);
  ctx->eng->set_stream((hipStream_t)s);
  return LD_OK;
}

int ld_rfft(ld_ctx* ctx, const void* u, void* fu, int nfields) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(u); LD_FIELD(fu);
  if (nfields < 1) { set_error("nfields must be >= 1"); return LD_ERR_ARG; }
  return ctx->eng->rfft(u, fu, nfields);
}

int ld_irfft(ld_ctx* ctx, const void* fu, void* u, int nfields) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(fu); LD_FIELD(u);
  if (nfields < 1) { set_error("nfields must be >= 1"); return LD_ERR_ARG; }
  return ctx->eng->irfft(fu, u, nfields);
}

int ld_forcing(ld_ctx* ctx, const void* uk, void* fk, double* ek_out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_FIELD(fk);
  return ctx->eng->forcing(uk, fk, ek_out);
}

int ld_clip_zero(ld_ctx* ctx, void* x, void* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(x); LD_FIELD(out);
  return ctx->eng->clip_zero(x, out);
}

int ld_full_rhs(ld_ctx* ctx, const void* uk, const void* n, const double* coord, int64_t np, void* ku, void* kn,
                double* kp, void* fc, double visc, double forc, void* u_out, void* v_out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_FIELD(n); LD_FIELD(ku); LD_FIELD(kn); LD_FIELD(fc);
  LD_ALIGNED(u_out); LD_ALIGNED(v_out);
  if (np < 0) { set_error("np must be >= 0"); return LD_ERR_ARG; }
  if (np > 0) { LD_NONNULL(coord); LD_NONNULL(kp); }
  return ctx->eng->full_rhs(uk, n, coord, np, ku, kn, kp, fc, visc, forc, u_out, v_out);
}

int ld_rk4(ld_ctx* ctx, double h, const void* uk, void* n, const double* coord, int64_t np, void* uknew, void* nnew,
           double* coord_new) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_FIELD(n); LD_FIELD(uknew); LD_FIELD(nnew);
  if (np < 0) { set_error("np must be >= 0"); return LD_ERR_ARG; }
  if (np > 0) { LD_NONNULL(coord); LD_NONNULL(coord_new); }
  return ctx->eng->rk4(h, uk, n, coord, np, uknew, nnew, coord_new);
}

int ld_state_set(ld_ctx* ctx, const void* uk, const void* n, const double* coord, int64_t np) {
  LD_CHECK_CTX(ctx); LD_FIELD(uk); LD_FIELD(n);
  if (np < 0) { set_error("np must be >= 0"); return LD_ERR_ARG; }
  if (np > 0) LD_NONNULL(coord);
  return ctx->eng->state_set(uk, n, coord, np);
}

int ld_state_get(ld_ctx* ctx, void* uk, void* n, double* coord) {
  LD_CHECK_CTX(ctx); LD_ALIGNED(uk); LD_ALIGNED(n);
  return ctx->eng->state_get(uk, n, coord);
}

int ld_advance(ld_ctx* ctx, double h, int32_t nsteps) {
  LD_CHECK_CTX(ctx);
  if (nsteps < 0) { set_error("nsteps must be >= 0"); return LD_ERR_ARG; }
  return ctx->eng->advance(h, nsteps);
}

int ld_advance_plain(ld_ctx* ctx, double h, int32_t nsteps) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx);
  if (nsteps < 0) { set_error("nsteps must be >= 0"); return LD_ERR_ARG; }
  return ctx->eng->advance_plain(h, nsteps);
}

int ld_step(ld_ctx* ctx, double h, void* uk, void* n, double* coord, int64_t np, int32_t nsteps) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_FIELD(n);
  if (np < 0 || nsteps < 0) { set_error("np and nsteps must be >= 0"); return LD_ERR_ARG; }
  if (np > 0) LD_NONNULL(coord);
  int rc = ctx->eng->state_set(uk, n, coord, np);
  if (rc != LD_OK) return rc;
  const int adv = ctx->eng->advance(h, nsteps);
  if (adv != LD_OK && adv != LD_ERR_BLOWUP) return adv;
  rc = ctx->eng->state_get(uk, n, coord);
  if (rc != LD_OK) return rc;
  return adv;
}

int ld_status(ld_ctx* ctx, double* last_max, int32_t* blown_up) {
  LD_CHECK_CTX(ctx);
  return ctx->eng->status(last_max, blown_up);
}

int ld_tuning(ld_ctx* ctx, int32_t* ylines, double* us) {
  LD_CHECK_CTX(ctx);
  return ctx->eng->tuning(ylines, us);
}

int ld_tuning_all(ld_ctx* ctx, int32_t* choice, double* us) {
  LD_CHECK_CTX(ctx);
  return ctx->eng->tuning_all(choice, us);
}

int ld_status_all(ld_ctx* ctx, double* last_max, int32_t* blown_up) {
  LD_CHECK_CTX(ctx);
  return ctx->eng->status_all(last_max, blown_up);
}

int ld_profile(ld_ctx* ctx, int32_t kernel, int32_t enable) {
  LD_CHECK_CTX(ctx);
  if (kernel < 0 || kernel > LD_K_DEPOSIT) { set_error("unknown kernel class %d", kernel); return LD_ERR_ARG; }
  return ctx->eng->profile(kernel, enable);
}

int ld_profile_read(ld_ctx* ctx, double* total_ms, int64_t* launches) {
  LD_CHECK_CTX(ctx); LD_NONNULL(total_ms); LD_NONNULL(launches);
  return ctx->eng->profile_read(total_ms, launches);
}

int ld_spectrum(ld_ctx* ctx, const void* f, int32_t nfields, double weight, double* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(f); LD_NONNULL(out);
  if (nfields < 1) { set_error("nfields must be >= 1"); return LD_ERR_ARG; }
  return ctx->eng->spectrum(f, nfields, weight, out);
}

int ld_spectrum_noalloc(ld_ctx* ctx, const void* f, int32_t nfields, double weight, double* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(f); LD_NONNULL(out);
  if (nfields < 1) { set_error("nfields must be >= 1"); return LD_ERR_ARG; }
  double* ws = g_scratch[ctx->device];
  if (!ws) { set_error("ld_spectrum_noalloc: no scratch address for device %d", ctx->device); return LD_ERR_HIP; }
  return ctx->eng->spectrum_ws(f, nfields, weight, out, ws, LD_SCRATCH_DOUBLES);
}

int ld_flux(ld_ctx* ctx, const void* uk, const void* ku, double* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_FIELD(ku); LD_NONNULL(out);
  return ctx->eng->flux(uk, ku, out);
}

int ld_dissipation(ld_ctx* ctx, const void* uk, double* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_NONNULL(out);
  return ctx->eng->dissipation(uk, out);
}

int ld_divergence(ld_ctx* ctx, const void* uk, void* div) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_FIELD(div);
  return ctx->eng->divergence(uk, div);
}

int ld_droplet_stats(ld_ctx* ctx, double* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_NONNULL(out);
  return ctx->eng->droplet_stats(out);
}

int ld_field_stats(ld_ctx* ctx, const void* x, int32_t nfields, double* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(x); LD_NONNULL(out);
  if (nfields < 1) { set_error("nfields must be >= 1"); return LD_ERR_ARG; }
  return ctx->eng->stats(x, nfields, out);
}

int ld_collide_pairs(ld_ctx* ctx, const double* rows, const double* radius, int64_t np, double rmax,
                     int64_t* pairs, int64_t max_pairs, int64_t* npairs) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_NONNULL(npairs);
  if (np < 0 || max_pairs < 0) { set_error("np and max_pairs must be >= 0"); return LD_ERR_ARG; }
  if (np > 1) { LD_NONNULL(rows); LD_NONNULL(radius); LD_NONNULL(pairs); }
  return ctx->eng->collide(rows, radius, np, rmax, pairs, max_pairs, npairs);
}

int ld_pair_stats(ld_ctx* ctx, const double* rows, int64_t np, double rmax, int32_t nbins, int64_t* counts,
                  double* wsum, double* wneg) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_NONNULL(counts); LD_NONNULL(wsum); LD_NONNULL(wneg);
  if (np < 0) { set_error("np must be >= 0"); return LD_ERR_ARG; }
  if (np > 1) LD_NONNULL(rows);
  return ctx->eng->pair_stats(rows, np, rmax, nbins, counts, wsum, wneg);
}

int ld_coalesce(ld_ctx* ctx, double* rows, int64_t np, const int64_t* pairs, int64_t npairs, int64_t* events,
                int64_t* nevents) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_NONNULL(nevents);
  if (np < 0 || npairs < 0) { set_error("np and npairs must be >= 0"); return LD_ERR_ARG; }
  if (npairs > 0) { LD_NONNULL(rows); LD_NONNULL(pairs); LD_NONNULL(events); }
  return ctx->eng->coalesce(rows, np, pairs, npairs, events, nevents);
}

int ld_interp_cosine(ld_ctx* ctx, const double* coord, int64_t np, const void* fields, int32_t ncomp, double* out) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx);
  if (np < 0 || ncomp < 1) { set_error("np must be >= 0 and ncomp >= 1"); return LD_ERR_ARG; }
  if (np > 0) { LD_NONNULL(coord); LD_FIELD(fields); LD_NONNULL(out); }
  if (np > 0x7fffffff) { set_error("np too large"); return LD_ERR_ARG; }
  return ctx->eng->interp(coord, np, fields, ncomp, out);
}

int ld_exterp_cosine_scalar(ld_ctx* ctx, const double* coord, int64_t np, const double* w, void* c) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(c);
  if (np < 0) { set_error("np must be >= 0"); return LD_ERR_ARG; }
  if (np > 0) { LD_NONNULL(coord); LD_NONNULL(w); }
  if (np > 0x7fffffff) { set_error("np too large"); return LD_ERR_ARG; }
  return ctx->eng->exterp(coord, np, w, c);
}

int ld_prhs(ld_ctx* ctx, const double* coord, int64_t np, const void* u, const void* us, const void* n, void* fc,
            double* rhs, double* interpmat, double* exterpmat) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(u); LD_FIELD(us); LD_FIELD(n); LD_FIELD(fc);
  if (np < 0) { set_error("np must be >= 0"); return LD_ERR_ARG; }
  if (np > 0) { LD_NONNULL(coord); LD_NONNULL(rhs); LD_NONNULL(interpmat); LD_NONNULL(exterpmat); }
  if (np > 0x7fffffff) { set_error("np too large"); return LD_ERR_ARG; }
  return ctx->eng->prhs(coord, np, u, us, n, fc, rhs, interpmat, exterpmat);
}

static bool pow2(int n) { return n >= 4 && n <= 4096 && (n & (n - 1)) == 0; }

int ld_spectral_resample(ld_ctx* ctx, const void* src, int32_t n_src, void* dst, int32_t nfields) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(src); LD_FIELD(dst);
  if (!pow2(n_src) || nfields < 1) { set_error("n_src must be a power of two >= 4 and nfields >= 1"); return LD_ERR_ARG; }
  return ctx->eng->resample(src, n_src, dst, nfields);
}

int ld_velocity_gradient_spec(ld_ctx* ctx, const void* uk, void* gk, double scale) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(uk); LD_FIELD(gk);
  return ctx->eng->grad_spec(uk, gk, scale);
}

int ld_fields_to_rows(ld_ctx* ctx, const void* f, int32_t nfields, float* rows, float scale) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(f); LD_NONNULL(rows);
  if (nfields < 1 || nfields > 32) { set_error("nfields must be in [1, 32]"); return LD_ERR_ARG; }
  return ctx->eng->fields_to_rows(f, nfields, rows, scale);
}

static int subgrid_ratio_ok(ld_ctx* ctx, int32_t r) {
  if (r < 1 || r > 4 || ctx->p.N % r != 0 || 3 * r * r * r > 192) {
    set_error("ratio must be in [1, 4] and divide N");
    return 0;
  }
  return 1;
}

int ld_subgrid_compose(ld_ctx* ctx, const float* net, int32_t ratio, const void* base, void* out, float scale) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_NONNULL(net); LD_FIELD(base); LD_FIELD(out);
  if (!subgrid_ratio_ok(ctx, ratio)) return LD_ERR_ARG;
  return ctx->eng->subgrid(net, nullptr, ratio, base, nullptr, out, scale);
}

int ld_subgrid_residual(ld_ctx* ctx, const void* fine, const void* base, int32_t ratio, float* net, float scale) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(fine); LD_FIELD(base); LD_NONNULL(net);
  if (!subgrid_ratio_ok(ctx, ratio)) return LD_ERR_ARG;
  return ctx->eng->subgrid(nullptr, net, ratio, base, fine, nullptr, scale);
}

int ld_droplets_frozen_rk4(ld_ctx* ctx, double* rows, int64_t np, const void* u, double h, int32_t nsteps) {
  LD_CHECK_CTX(ctx); LD_ONE(ctx); LD_FIELD(u);
  if (np < 0 || nsteps < 0) { set_error("np and nsteps must be >= 0"); return LD_ERR_ARG; }
  if (np > 0) LD_NONNULL(rows);
  if (np > 0x1fffffff) { set_error("np too large (at most 2^29 - 1)"); return LD_ERR_ARG; }
  return ctx->eng->frozen_rk4(rows, np, u, h, nsteps);
}

}  // extern "C"
