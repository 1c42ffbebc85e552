// Engine<T, N>: buffer plan and launch sequence of the step for one
// precision and grid size.  Instantiated once per (T, N) in ld_inst.hip.
#pragma once
#include "ld_engine_base.hpp"
#include "ld_particles.hpp"
#include <cmath>
#include <utility>
#include <vector>

namespace ld {

#define LD_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_error("%s failed: %s", #call, hipGetErrorString(e_));                        \
      return LD_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)
#define LD_LAUNCHED()                                                                  \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess) {                                                            \
      set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
      return LD_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)
#define LD_TRY(x)                 \
  do {                            \
    int rc_ = (x);                \
    if (rc_ != LD_OK) return rc_; \
  } while (0)

template <typename T> struct PrecTraits;
// Lines per block of a strided pass: 16 complex64 or 8 complex128 = 128 B
// contiguous per element row.
template <> struct PrecTraits<float> { static constexpr int M = 16; using PT = float; };
template <> struct PrecTraits<double> { static constexpr int M = 8; using PT = double; };

// z-pass tile height P for F fields: <= 256 threads, <= 64 KB LDS, >= 2.
template <typename T, int N, int F> constexpr int zTileP() {
  constexpr int N1 = Plan<N>::N1, Nf = N / 2 + 1;
  int P = 2;
  for (;;) {
    const int P2 = 2 * P;
    const int threads = F * (P2 / 2) * N1;
    const long lds = (long)F * (P2 / 2) * 2 * Nf * (long)sizeof(C<T>);
    if (threads > 256 || lds > 65536 || P2 > N * N) break;
    P = P2;
  }
  return P;
}

template <typename T, int N>
struct Engine final : EngineBase {
  using PT = typename PrecTraits<T>::PT;
  static constexpr int M = PrecTraits<T>::M;
  static constexpr int N1 = Plan<N>::N1, Nf = N / 2 + 1;
  static constexpr size_t CF = Geo<N>::C, RF = Geo<N>::R;
  static constexpr int P1 = zTileP<T, N, 1>(), P3 = zTileP<T, N, 3>(), P6 = zTileP<T, N, 6>();
  static constexpr int NBLK_Z1 = N * N / P1;       // blocks of a 1-field z pass
  static constexpr int NBLK_X = N * Nf / M;         // blocks of an x pass
  static constexpr int NBLK_SUM = 1024;             // blocks of k_sum2

  ld_params p;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  C<T>* tw = nullptr;
  C<T>*uk = nullptr, *us = nullptr, *acc = nullptr, *w6 = nullptr;
  T *u = nullptr, *v = nullptr, *n_state = nullptr, *n_stage = nullptr, *nacc = nullptr, *fc = nullptr;
  double *partials = nullptr, *clip_stage = nullptr, *clip_end = nullptr, *clip_none = nullptr;
  double *maxpart = nullptr, *ek = nullptr, *fac = nullptr, *lastmax = nullptr;
  int *flag = nullptr, *shells = nullptr;
  int smax = 0;
  int64_t cap = 0, np = 0;
  double *x0 = nullptr, *sx = nullptr, *kxp = nullptr;
  PT *v0 = nullptr, *sv = nullptr, *kvp = nullptr;
  double *rows_a = nullptr, *rows_b = nullptr, *wrow = nullptr;
  // spatial sort: slot -> input row, bin keys/ranks, bin starts, double buffers
  int64_t *rid = nullptr, *rid2 = nullptr;
  int *pkey = nullptr, *prank = nullptr, *hist = nullptr, *bsum = nullptr;
  double* x0b = nullptr;
  PT* v0b = nullptr;
  bool sort_enabled = true;
  static constexpr int NB = Bins<N>::NB;
  static constexpr int NBS = (NB + 4095) / 4096;
  DropPhys dp;

  explicit Engine(const ld_params& prm) : p(prm) {}
  ~Engine() override {
    for (void* a : allocs) (void)hipFree(a);
  }

  template <class X> int alloc(X*& ptr, size_t count) {
    void* q = nullptr;
    if (count == 0) count = 1;
    LD_HIP(hipMalloc(&q, count * sizeof(X)));
    allocs.push_back(q);
    ptr = (X*)q;
    return LD_OK;
  }

  int init() override {
    if (p.lp != 1) { set_error("only lp = 1 (Laplacian viscosity) is supported"); return LD_ERR_ARG; }
    if (p.nshells < 0 || p.nshells > 8) { set_error("nshells must be in [0, 8]"); return LD_ERR_ARG; }
    for (int i = 0; i < p.nshells; ++i) {
      if (p.shell_no[i] < 1 || p.shell_no[i] > 8 || p.shell_no[i] > N / 2 - 1) {
        set_error("forced shell %d out of range [1, min(8, N/2-1)]", p.shell_no[i]);
        return LD_ERR_ARG;
      }
      smax = p.shell_no[i] > smax ? p.shell_no[i] : smax;
    }
    if (p.max_particles < 0 || p.max_particles > (int64_t)0x7fffffff) { set_error("max_particles out of range"); return LD_ERR_ARG; }
    cap = p.max_particles;
    // twiddles exp(-2 pi i m / N), rounded from long double
    std::vector<C<T>> h(N);
    for (int m = 0; m < N; ++m) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)N;
      h[m] = mkc<T>((T)cosl(ang), (T)sinl(ang));
    }
    LD_TRY(alloc(tw, N));
    LD_HIP(hipMemcpy(tw, h.data(), N * sizeof(C<T>), hipMemcpyHostToDevice));
    LD_TRY(alloc(uk, 3 * CF)); LD_TRY(alloc(us, 3 * CF)); LD_TRY(alloc(acc, 3 * CF)); LD_TRY(alloc(w6, 6 * CF));
    LD_TRY(alloc(u, 3 * RF)); LD_TRY(alloc(v, 3 * RF));
    LD_TRY(alloc(n_state, RF)); LD_TRY(alloc(n_stage, RF)); LD_TRY(alloc(nacc, RF)); LD_TRY(alloc(fc, RF));
    const int npart = NBLK_Z1 > NBLK_SUM ? NBLK_Z1 : NBLK_SUM;
    LD_TRY(alloc(partials, 2 * (size_t)npart));
    LD_TRY(alloc(clip_stage, 4)); LD_TRY(alloc(clip_end, 4)); LD_TRY(alloc(clip_none, 4));
    LD_TRY(alloc(maxpart, NBLK_X)); LD_TRY(alloc(ek, Nf)); LD_TRY(alloc(fac, Nf)); LD_TRY(alloc(lastmax, 1));
    LD_TRY(alloc(flag, 1)); LD_TRY(alloc(shells, 8));
    LD_HIP(hipMemset(fac, 0, Nf * sizeof(double)));
    LD_HIP(hipMemset(ek, 0, Nf * sizeof(double)));
    LD_HIP(hipMemset(flag, 0, sizeof(int)));
    LD_HIP(hipMemcpy(shells, p.shell_no, 8 * sizeof(int), hipMemcpyHostToDevice));
    const double cn[4] = {1.0, 1.0, 0.0, 0.0};
    LD_HIP(hipMemcpy(clip_none, cn, sizeof(cn), hipMemcpyHostToDevice));
    LD_HIP(hipMemcpy(clip_stage, cn, sizeof(cn), hipMemcpyHostToDevice));
    LD_HIP(hipMemcpy(clip_end, cn, sizeof(cn), hipMemcpyHostToDevice));
    if (cap > 0) {
      LD_TRY(alloc(x0, 3 * cap)); LD_TRY(alloc(sx, 3 * cap)); LD_TRY(alloc(kxp, 3 * cap));
      LD_TRY(alloc(v0, 4 * cap)); LD_TRY(alloc(sv, 4 * cap)); LD_TRY(alloc(kvp, 4 * cap));
      LD_TRY(alloc(rows_a, 7 * cap)); LD_TRY(alloc(rows_b, 7 * cap)); LD_TRY(alloc(wrow, cap));
      LD_TRY(alloc(rid, cap)); LD_TRY(alloc(rid2, cap)); LD_TRY(alloc(pkey, cap)); LD_TRY(alloc(prank, cap));
      LD_TRY(alloc(hist, NB)); LD_TRY(alloc(bsum, NBS));
      LD_TRY(alloc(x0b, 3 * cap)); LD_TRY(alloc(v0b, 4 * cap));
    }
    static_assert(NBS <= 4096, "bin-start scan supports at most 4096 partial sums");
    dp.L = p.L;
    dp.dx = p.L / N;                       // np.linspace(0, L, N, endpoint=False) step
    dp.pi_2dx = M_PI / (2.0 * dp.dx);
    dp.inv_vol = 1.0 / (dp.dx * dp.dx * dp.dx);
    dp.factor = p.factor; dp.growth = p.growthfactor; dp.st_s = p.st_s; dp.tau_eta = p.tau_eta; dp.g = p.g;
    return LD_OK;
  }

  void set_stream(hipStream_t s) override { st = s; }

  // ------------------------------------------------------------ launch helpers
  template <int AXIS, int DIR> int pass(const C<T>* in, C<T>* out, int nf, T scale) {
    hipLaunchKernelGGL((k_pass<T, N, M, AXIS, DIR>), dim3(NBLK_X, nf), dim3(M * N1), 0, st, in, out, CF,
                       (const C<T>*)tw, scale);
    LD_LAUNCHED();
    return LD_OK;
  }
  template <int P, int FIN, int FOUT, class Op> int zrun(const Op& op) {
    hipLaunchKernelGGL((k_z<T, N, P, FIN, FOUT, Op>), dim3(N * N / P), dim3(ZThreads<N, P, FIN, FOUT>::v), 0, st, op,
                       (const C<T>*)tw);
    LD_LAUNCHED();
    return LD_OK;
  }
  template <class X> int fill(X* x, size_t n, X val) {
    hipLaunchKernelGGL((k_fill<X>), dim3(1024), dim3(256), 0, st, x, n, val);
    LD_LAUNCHED();
    return LD_OK;
  }
  int shell(const C<T>* uin, double fscale) {
    const double norm0 = 1.0 / ((double)N * N * N * N * N * N / std::pow(2.0 * M_PI, 3.0));
    if (p.nshells == 0) return LD_OK;
    hipLaunchKernelGGL((k_shell<T, N>), dim3(1), dim3(256), 0, st, uin, CF, smax, (const int*)shells, p.nshells,
                       p.f0, fscale, norm0, ek, fac);
    LD_LAUNCHED();
    return LD_OK;
  }
  int clip_sums(const T* x, double* clip) {
    hipLaunchKernelGGL((k_sum2<T>), dim3(NBLK_SUM), dim3(256), 0, st, x, RF, partials);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_clip_final, dim3(1), dim3(256), 0, st, (const double*)partials, NBLK_SUM,
                       1.0 / ((double)N * N * N), clip);
    LD_LAUNCHED();
    return LD_OK;
  }
  static T inv() { return (T)(1.0 / N); }

  // ------------------------------------------------------------ transforms
  int rfft(const void* uin, void* fuout, int nf) override {
    const T* src = (const T*)uin;
    C<T>* dst = (C<T>*)fuout;
    for (int f = 0; f < nf; ++f) {
      OpZFwd<T, N, P1, 1> op{src + f * RF, dst + f * CF, RF, CF};
      LD_TRY((zrun<P1, 0, 1>(op)));
    }
    LD_TRY((pass<1, -1>(dst, dst, nf, (T)1)));
    LD_TRY((pass<0, -1>(dst, dst, nf, (T)1)));
    return LD_OK;
  }
  int irfft(const void* fuin, void* uout, int nf) override {
    const C<T>* src = (const C<T>*)fuin;
    T* dst = (T*)uout;
    for (int f0 = 0; f0 < nf; f0 += 6) {
      const int k = nf - f0 < 6 ? nf - f0 : 6;
      LD_TRY((pass<0, +1>(src + f0 * CF, w6, k, inv())));
      LD_TRY((pass<1, +1>(w6, w6, k, inv())));
      for (int f = 0; f < k; ++f) {
        OpZInv<T, N, P1, 1> op{w6 + f * CF, dst + (f0 + f) * RF, CF, RF, inv()};
        LD_TRY((zrun<P1, 1, 0>(op)));
      }
    }
    return LD_OK;
  }

  int forcing(const void* ukin, void* fkout, double* ek_out) override {
    LD_TRY(shell((const C<T>*)ukin, p.isforcing));
    const size_t nb = (CF + 255) / 256;
    hipLaunchKernelGGL((k_forcing_apply<T, N>), dim3(nb), dim3(256), 0, st, (const C<T>*)ukin, (C<T>*)fkout, CF,
                       (const double*)fac);
    LD_LAUNCHED();
    if (ek_out) LD_HIP(hipMemcpyAsync(ek_out, ek, Nf * sizeof(double), hipMemcpyDeviceToDevice, st));
    return LD_OK;
  }

  int clip_zero(void* x, void* out) override {
    LD_TRY(clip_sums((const T*)x, clip_end));
    hipLaunchKernelGGL((k_clip_apply<T>), dim3(1024), dim3(256), 0, st, (T*)x, (T*)out, RF, (const double*)clip_end, 1);
    LD_LAUNCHED();
    return LD_OK;
  }

  // ------------------------------------------------------------ stage pieces
  // u, omega -> physical u (to uphys) and the z/y-forward of u x omega in w6[0..3).
  int nonlinear(const C<T>* usp, T* uphys) {
    hipLaunchKernelGGL((k_x_uw<T, N, M>), dim3(NBLK_X), dim3(M * N1), 0, st, usp, w6, CF, (const C<T>*)tw, inv());
    LD_LAUNCHED();
    LD_TRY((pass<1, +1>(w6, w6, 6, inv())));
    OpZCross<T, N, P6> oc{w6, CF, uphys, inv()};
    LD_TRY((zrun<P6, 6, 3>(oc)));
    LD_TRY((pass<1, -1>(w6, w6, 3, (T)1)));
    return LD_OK;
  }
  // v_hat (x-inverse, in w6[0..3)) -> physical v and z/y... up to the x-inverse
  // of v (leaves nv forward-z in w6[0..3)).
  int velocity_nv(T* vphys, const T* ns, const double* clip) {
    LD_TRY((pass<1, +1>(w6, w6, 3, inv())));
    OpZVel<T, N, P3> ov{w6, CF, vphys, ns, clip, inv(), (T)(p.tps * p.g)};
    LD_TRY((zrun<P3, 3, 3>(ov)));
    return LD_OK;
  }
  // fc -> z forward into w6[3]; y forward of w6[0..4); kn_hat x-inverse in w6[0]; y-inverse.
  int density_rhs(const T* fcin) {
    OpZFwd<T, N, P1, 1> of{fcin, w6 + 3 * CF, RF, CF};
    LD_TRY((zrun<P1, 0, 1>(of)));
    LD_TRY((pass<1, -1>(w6, w6, 4, (T)1)));
    hipLaunchKernelGGL((k_x_kn<T, N, M>), dim3(NBLK_X), dim3(M * N1), 0, st, w6, CF, (const C<T>*)tw, inv());
    LD_LAUNCHED();
    LD_TRY((pass<1, +1>(w6, w6, 1, inv())));
    return LD_OK;
  }

  // One RK4 stage s (0..3) of the fused step on the context state.
  int stage(int s, double h) {
    static const double a_coef[5] = {0.0, 0.5, 0.5, 1.0, 0.0};
    static const double w_coef[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const C<T>* usp = (s == 0) ? uk : us;
    const T* ns = (s == 0) ? n_state : n_stage;
    const double wh = (s == 0 || s == 3) ? h / 6.0 : h / 3.0;   // forced-dns-sm.py:342-358
    const double anext = a_coef[s + 1] * h;
    (void)w_coef;
    LD_TRY(shell(usp, p.isforcing));
    LD_TRY(nonlinear(usp, u));
    RhsArgs<T> ra;
    ra.nl = w6; ra.us = usp; ra.ubase = uk; ra.acc = acc; ra.usnext = (s < 3) ? us : nullptr; ra.ku = nullptr;
    ra.fac = fac; ra.fs = CF; ra.nu_eff = (T)(p.nu * p.isexplicit); ra.tps = (T)p.tps;
    ra.wacc = (T)wh; ra.anext = (T)anext; ra.scale = inv(); ra.acc_init = (s == 0);
    hipLaunchKernelGGL((k_x_rhs<T, N, M>), dim3(NBLK_X), dim3(M * N1), 0, st, ra, (const C<T>*)tw);
    LD_LAUNCHED();
    LD_TRY(velocity_nv(v, ns, clip_stage));
    LD_TRY(fill<T>(fc, RF, (T)0));
    if (np > 0) {
      PStage<T, PT> ps;
      ps.np = (int)np; ps.cap = (int)cap; ps.stage = s;
      ps.x0 = x0; ps.v0 = v0; ps.sx = sx; ps.sv = sv; ps.kx = kxp; ps.kv = kvp;
      ps.u = u; ps.v = v; ps.n = ns; ps.clip = clip_stage; ps.fc = fc;
      ps.ah = a_coef[s] * h; ps.wh = wh; ps.P = dp;
      // 4 lanes per droplet; block count a multiple of 8 for the XCD mapping
      const unsigned nb = (unsigned)(((4 * np + 255) / 256 + 7) / 8 * 8);
      hipLaunchKernelGGL((k_particles_stage<T, PT, N>), dim3(nb), dim3(256), 0, st, ps);
      LD_LAUNCHED();
    }
    LD_TRY(density_rhs(fc));
    OpZKn<T, N, P1> ok;
    ok.w = w6; ok.fs = CF; ok.scale = inv(); ok.kn_out = nullptr;
    ok.n0 = n_state; ok.nacc = nacc; ok.nnext = (s < 3) ? n_stage : nullptr;
    ok.wacc = (T)wh; ok.anext = (T)anext; ok.acc_init = (s == 0);
    ok.sum_target = (s < 3) ? 0 : 1; ok.partials = partials;
    LD_TRY((zrun<P1, 1, 0>(ok)));
    hipLaunchKernelGGL(k_clip_final, dim3(1), dim3(256), 0, st, (const double*)partials, NBLK_Z1,
                       1.0 / ((double)N * N * N), (s < 3) ? clip_stage : clip_end);
    LD_LAUNCHED();
    return LD_OK;
  }

  // RK4 over the context state: acc <- uknew, nacc <- nnew, droplet base <- sump.
  int rk4_stages(double h) {
    LD_TRY(clip_sums(n_state, clip_stage));   // stage-1 clip_zero(n) (forced-dns-sm.py:341)
    for (int s = 0; s < 4; ++s) LD_TRY(stage(s, h));
    return LD_OK;
  }

  // Step epilogue (forced-dns-sm.py:583-620).
  int epilogue() {
    hipLaunchKernelGGL((k_clip_apply<T>), dim3(1024), dim3(256), 0, st, nacc, n_state, RF, (const double*)clip_end, 1);
    LD_LAUNCHED();
    LD_TRY((pass<0, +1>(acc, w6, 3, inv())));
    LD_TRY((pass<1, +1>(w6, w6, 3, inv())));
    OpZRound<T, N, P3> orr{w6, CF, inv()};
    LD_TRY((zrun<P3, 3, 3>(orr)));
    LD_TRY((pass<1, -1>(w6, w6, 3, (T)1)));
    hipLaunchKernelGGL((k_x_proj<T, N, M>), dim3(NBLK_X), dim3(M * N1), 0, st, (const C<T>*)w6, uk, CF,
                       (const C<T>*)tw, maxpart);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, st, (const double*)maxpart, NBLK_X,
                       100.0 * (double)N * N * N, flag, lastmax);
    LD_LAUNCHED();
    return LD_OK;
  }

  // ------------------------------------------------------------ state
  int set_rows(const double* coord, int64_t n_) {
    if (n_ > cap) { set_error("%lld droplets exceed the context capacity %lld", (long long)n_, (long long)cap); return LD_ERR_CAPACITY; }
    np = n_;
    if (np > 0) {
      hipLaunchKernelGGL((k_rows_to_soa<PT>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, coord, (int)np,
                         (int)cap, x0, v0);
      LD_LAUNCHED();
      hipLaunchKernelGGL(k_iota64, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, rid, (int)np);
      LD_LAUNCHED();
    }
    return LD_OK;
  }
  int get_rows(double* coord) {
    if (np > 0) {
      hipLaunchKernelGGL((k_soa_to_rows<PT>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                         (const double*)x0, (const PT*)v0, (const int64_t*)rid, (int)np, (int)cap, coord);
      LD_LAUNCHED();
    }
    return LD_OK;
  }
  int state_set(const void* ukin, const void* nin, const double* coord, int64_t n_) override {
    LD_HIP(hipMemcpyAsync(uk, ukin, 3 * CF * sizeof(C<T>), hipMemcpyDeviceToDevice, st));
    LD_HIP(hipMemcpyAsync(n_state, nin, RF * sizeof(T), hipMemcpyDeviceToDevice, st));
    return set_rows(coord, n_);
  }
  int state_get(void* ukout, void* nout, double* coord) override {
    if (ukout) LD_HIP(hipMemcpyAsync(ukout, uk, 3 * CF * sizeof(C<T>), hipMemcpyDeviceToDevice, st));
    if (nout) LD_HIP(hipMemcpyAsync(nout, n_state, RF * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (coord) LD_TRY(get_rows(coord));
    return LD_OK;
  }
  // Counting sort of the droplet slots by spatial bin of the base position.
  int sort_droplets() {
    if (np < 2 || !sort_enabled) return LD_OK;
    const unsigned pb = (unsigned)((np + 255) / 256);
    LD_TRY(fill<int>(hist, NB, 0));
    hipLaunchKernelGGL((k_bin_count<N>), dim3(pb), dim3(256), 0, st, (const double*)x0, (int)np, (int)cap, p.L,
                       dp.dx, pkey, prank, hist);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_local, dim3(NBS), dim3(1024), 0, st, hist, NB, bsum);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(1024), 0, st, bsum, NBS);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_add, dim3((NB + 255) / 256), dim3(256), 0, st, hist, NB, (const int*)bsum);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_bin_scatter<PT>), dim3(pb), dim3(256), 0, st, (int)np, (int)cap, (const int*)pkey,
                       (const int*)prank, (const int*)hist, (const double*)x0, (const PT*)v0, (const int64_t*)rid,
                       x0b, v0b, rid2);
    LD_LAUNCHED();
    std::swap(x0, x0b);
    std::swap(v0, v0b);
    std::swap(rid, rid2);
    return LD_OK;
  }

  int advance(double h, int nsteps) override {
    for (int i = 0; i < nsteps; ++i) {
      LD_TRY(sort_droplets());
      LD_TRY(rk4_stages(h));
      LD_TRY(epilogue());
    }
    int32_t blown = 0;
    LD_HIP(hipMemcpyAsync(&blown, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    LD_HIP(hipStreamSynchronize(st));
    if (blown) { set_error("Threshold exceeded: max Re(u_hat) > 100 N^3"); return LD_ERR_BLOWUP; }
    return LD_OK;
  }
  int status(double* last_max, int32_t* blown) override {
    if (last_max) LD_HIP(hipMemcpyAsync(last_max, lastmax, sizeof(double), hipMemcpyDeviceToHost, st));
    if (blown) LD_HIP(hipMemcpyAsync(blown, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    LD_HIP(hipStreamSynchronize(st));
    return LD_OK;
  }

  int rk4(double h, const void* ukin, void* nio, const double* coord, int64_t n_, void* uknew, void* nnew,
          double* coord_new) override {
    LD_TRY(state_set(ukin, nio, coord, n_));
    LD_TRY(rk4_stages(h));
    LD_HIP(hipMemcpyAsync(uknew, acc, 3 * CF * sizeof(C<T>), hipMemcpyDeviceToDevice, st));
    LD_HIP(hipMemcpyAsync(nnew, nacc, RF * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (coord_new) LD_TRY(get_rows(coord_new));
    hipLaunchKernelGGL((k_relu<T>), dim3(1024), dim3(256), 0, st, (T*)nio, RF);   // in-place clip of n (:282 via :341)
    LD_LAUNCHED();
    return LD_OK;
  }

  int full_rhs(const void* ukin, const void* nin, const double* coord, int64_t n_, void* ku, void* kn, double* kp,
               void* fcout, double visc, double forc, void* u_out, void* v_out) override {
    if (n_ > cap) { set_error("%lld droplets exceed the context capacity %lld", (long long)n_, (long long)cap); return LD_ERR_CAPACITY; }
    const C<T>* ukp = (const C<T>*)ukin;
    T* uphys = u_out ? (T*)u_out : u;
    T* vphys = v_out ? (T*)v_out : v;
    T* fcp = (T*)fcout;
    LD_TRY(shell(ukp, p.isforcing * forc));
    LD_TRY(nonlinear(ukp, uphys));
    RhsArgs<T> ra;
    ra.nl = w6; ra.us = ukp; ra.ubase = ukp; ra.acc = nullptr; ra.usnext = nullptr; ra.ku = (C<T>*)ku;
    ra.fac = fac; ra.fs = CF; ra.nu_eff = (T)(p.nu * p.isexplicit * visc); ra.tps = (T)p.tps;
    ra.wacc = (T)0; ra.anext = (T)0; ra.scale = inv(); ra.acc_init = 0;
    hipLaunchKernelGGL((k_x_rhs<T, N, M>), dim3(NBLK_X), dim3(M * N1), 0, st, ra, (const C<T>*)tw);
    LD_LAUNCHED();
    // the physical-space v the droplets see uses visc as well (forced-dns-sm.py:319-321)
    LD_TRY(velocity_nv(vphys, (const T*)nin, clip_none));
    LD_TRY(fill<T>(fcp, RF, (T)0));
    if (n_ > 0) {
      hipLaunchKernelGGL((k_prhs<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_,
                         (const T*)uphys, (const T*)vphys, (const T*)nin, (const double*)clip_none, fcp, kp, rows_a,
                         wrow, dp);
      LD_LAUNCHED();
    }
    LD_TRY(density_rhs(fcp));
    OpZKn<T, N, P1> ok;
    ok.w = w6; ok.fs = CF; ok.scale = inv(); ok.kn_out = (T*)kn;
    ok.n0 = nullptr; ok.nacc = nullptr; ok.nnext = nullptr; ok.wacc = (T)0; ok.anext = (T)0; ok.acc_init = 0;
    ok.sum_target = -1; ok.partials = partials;
    LD_TRY((zrun<P1, 1, 0>(ok)));
    return LD_OK;
  }

  // ------------------------------------------------------------ droplet API
  int interp(const double* coord, int64_t n_, const void* f, int nc, double* out) override {
    if (n_ <= 0) return LD_OK;
    hipLaunchKernelGGL((k_interp<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_,
                       (const T*)f, nc, out, dp);
    LD_LAUNCHED();
    return LD_OK;
  }
  int exterp(const double* coord, int64_t n_, const double* w, void* c) override {
    LD_TRY(fill<T>((T*)c, RF, (T)0));
    if (n_ <= 0) return LD_OK;
    hipLaunchKernelGGL((k_exterp<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_, w,
                       (T*)c, dp);
    LD_LAUNCHED();
    return LD_OK;
  }
  int prhs(const double* coord, int64_t n_, const void* uf, const void* usf, const void* nf, void* fcout,
           double* rhs, double* interpm, double* exterpm) override {
    LD_TRY(fill<T>((T*)fcout, RF, (T)0));
    if (n_ <= 0) return LD_OK;
    hipLaunchKernelGGL((k_prhs<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_,
                       (const T*)uf, (const T*)usf, (const T*)nf, (const double*)clip_none, (T*)fcout, rhs, interpm,
                       exterpm, dp);
    LD_LAUNCHED();
    return LD_OK;
  }
};

}  // namespace ld
