// Engine interface: one implementation per (precision, N), chosen at ld_create.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../include/ld.h"

namespace ld {

void set_error(const char* fmt, ...);

struct EngineBase {
  virtual ~EngineBase() {}
  virtual int init() = 0;
  virtual void set_stream(hipStream_t s) = 0;
  virtual int rfft(const void* u, void* fu, int nf) = 0;
  virtual int irfft(const void* fu, void* u, int nf) = 0;
  virtual int forcing(const void* uk, void* fk, double* ek_out) = 0;
  virtual int clip_zero(void* x, void* out) = 0;
  virtual int full_rhs(const void* uk, const void* n, const double* coord, int64_t np, void* ku, void* kn,
                       double* kp, void* fc, double visc, double forc, void* u_out, void* v_out) = 0;
  virtual int rk4(double h, const void* uk, void* n, const double* coord, int64_t np, void* uknew, void* nnew,
                  double* coord_new) = 0;
  virtual int state_set(const void* uk, const void* n, const double* coord, int64_t np) = 0;
  virtual int state_get(void* uk, void* n, double* coord) = 0;
  virtual int advance(double h, int nsteps) = 0;
  virtual int advance_plain(double h, int nsteps) = 0;
  virtual int status(double* last_max, int32_t* blown) = 0;
  virtual int status_all(double* last_max, int32_t* blown) = 0;
  virtual int batch() const = 0;
  virtual int tuning(int32_t* ylines, double* us) = 0;
  virtual int tuning_all(int32_t* choice, double* us) = 0;
  virtual int interp(const double* coord, int64_t np, const void* f, int nc, double* out) = 0;
  virtual int exterp(const double* coord, int64_t np, const double* w, void* c) = 0;
  virtual int collide(const double* rows, const double* rad, int64_t np, double rmax, int64_t* pairs,
                      int64_t max_pairs, int64_t* count) = 0;
  virtual int pair_stats(const double* rows, int64_t np, double rmax, int nbins, int64_t* counts, double* wsum,
                         double* wneg) = 0;
  virtual int coalesce(double* rows, int64_t np, const int64_t* pairs, int64_t npairs, int64_t* events,
                       int64_t* nevents) = 0;
  virtual int spectrum(const void* f, int nf, double w, double* out) = 0;
  // spectrum with caller-provided scratch ws (ws_len doubles); allocates nothing
  virtual int spectrum_ws(const void* f, int nf, double w, double* out, double* ws, size_t ws_len) = 0;
  virtual int flux(const void* uk, const void* ku, double* out) = 0;
  virtual int dissipation(const void* uk, double* out) = 0;
  virtual int divergence(const void* uk, void* div) = 0;
  virtual int droplet_stats(double* out) = 0;
  virtual int stats(const void* x, int nf, double* out) = 0;
  virtual int profile(int kernel, int enable) = 0;
  virtual int profile_read(double* total_ms, int64_t* launches) = 0;
  virtual int prhs(const double* coord, int64_t np, const void* u, const void* us, const void* n, void* fc,
                   double* rhs, double* interp, double* exterp) = 0;
  virtual int resample(const void* src, int ns, void* dst, int nf) = 0;
  virtual int fields_to_rows(const void* f, int nc, float* rows, float s) = 0;
  virtual int grad_spec(const void* uk, void* gk, double s) = 0;
  virtual int subgrid(const float* net_in, float* net_out, int r, const void* base, const void* fine, void* out,
                      float s) = 0;
  virtual int frozen_rk4(double* rows, int64_t np, const void* u, double h, int nsteps) = 0;
};

#define LD_DECL_FACTORY(TN, NN) EngineBase* make_engine_##TN##_##NN(const ld_params& p);
LD_DECL_FACTORY(f32, 16) LD_DECL_FACTORY(f32, 32) LD_DECL_FACTORY(f32, 64)
LD_DECL_FACTORY(f32, 128) LD_DECL_FACTORY(f32, 256) LD_DECL_FACTORY(f32, 512)
LD_DECL_FACTORY(f64, 16) LD_DECL_FACTORY(f64, 32) LD_DECL_FACTORY(f64, 64)
LD_DECL_FACTORY(f64, 128) LD_DECL_FACTORY(f64, 256) LD_DECL_FACTORY(f64, 512)
#undef LD_DECL_FACTORY

}  // namespace ld
