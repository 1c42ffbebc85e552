// Engine<T, N>: buffer plan and launch sequence of the step for one
// precision and grid size.  Instantiated once per (T, N) in ld_inst.hip.
#pragma once
#include "ld_engine_base.hpp"
#include "ld_particles.hpp"
#include "ld_plain.hpp"
#include "ld_collide.hpp"
#include "ld_diag.hpp"
#include "ld_learned.hpp"
#include <algorithm>
#include <cmath>
#include <cstdlib>
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

// z-pass tile height P for F fields: <= 256 threads and <= 64 KB LDS (one
// field: <= 512 threads and <= 72 KB, measured at 256^3 FP32: 64 lines instead
// of 32 took 5.82 -> 5.79 ms per step), >= 2.
template <typename T, int N, int F> constexpr int zTileP() {
  constexpr int N1 = Plan<N>::N1, Nf = N / 2 + 1;
  constexpr int TMAX = F == 1 ? 512 : 256;
  constexpr long LMAX = F == 1 ? 73728 : 65536;
  int P = 2;
  for (;;) {
    const int P2 = 2 * P;
    const int threads = F * (P2 / 2) * N1;
    const long lds = (long)F * (P2 / 2) * 2 * Nf * (long)sizeof(C<T>);
    if (threads > TMAX || lds > LMAX || P2 > N * N) break;
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
  static constexpr int M3 = XM3<T, N, M>::v;         // lines per block of a field-parallel x pass
  static constexpr int NBLK_X3 = N * Nf / M3;
  static constexpr int M4 = XM4<T, N, M>::v;         // lines per block of the four-field x pass
  static constexpr int NBLK_X4 = N * Nf / M4;
  // the autotuner's alternatives: half the z-tile height (>= 2) and half the
  // x-pass lines per block (>= 2); a block's lines do not change any line's
  // arithmetic, so every choice gives the same bits
  static constexpr int P6H = P6 / 2 >= 2 ? P6 / 2 : P6, P3H = P3 / 2 >= 2 ? P3 / 2 : P3;
  // and twice the z-tile height where the block stays within 512 threads and 64 KB of LDS
  static constexpr bool zfits(int P, int F) {
    return F * (P / 2) * N1 <= 512 && (long)F * (P / 2) * 2 * Nf * (long)sizeof(C<T>) <= 65536 && P <= N * N;
  }
  static constexpr int P6D = zfits(2 * P6, 6) ? 2 * P6 : P6, P3D = zfits(2 * P3, 3) ? 2 * P3 : P3;
  static constexpr int M3H = M3 / 2 >= 2 ? M3 / 2 : M3, M4H = M4 / 2 >= 2 ? M4 / 2 : M4;
  // launch-geometry choices: 0 = default, 1 = the alternative above (2 = the
  // doubled z tile, z passes only), for the
  // u x omega z pass, the v z pass, the u/omega x pass, the RHS x pass and the
  // density x pass (set by the autotuner or LD_TUNE_FORCE=a,b,c,d,e)
  int vsel[5] = {0, 0, 0, 0, 0};
  double vus[5][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  static constexpr int NBLK_SUM = 1024;             // blocks of k_sum2

  ld_params p;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  C<T>* tw = nullptr;
  hipStream_t st2 = nullptr;        // density/droplet stream (created here)
  bool serial = false;              // no overlap of the flow and density chains
  unsigned pstage_lds = 0;          // dynamic LDS bytes per block of the droplet stage kernel
  // queue a stage's density chain after the next flow x pass (measured at 256^3:
  // 8.18 vs 8.31 ms/step; the 64^3 ensemble is faster in the plain order)
  bool late_density = N >= 128;
  // Dealiasing truncation: every mode with kz > N/3 is outside the mask
  // (kint >= kz), so the nonlinear term there is discarded and v_hat is zero.
  // The in-place y passes skip the kz > N/3 lines (kc = N/3; kc = N/2 runs
  // every line) where they carry the nonlinear term (x_rhs3 does not read
  // those lines) or v_hat (zero lines, already zero after the x pass).
  int kc = N / 3;
  // largest ky^2 + kz^2 of a (ky, kz) line holding a mode inside the mask
  // (round|k| <= kc <=> |k|^2 <= kc (kc + 1)); -1: no per-entry skipping
  int sphere() const { return kc < N / 2 ? kc * (kc + 1) : -1; }
  // u_hat of the context state has no nonzero mode outside the mask (checked
  // when the state is set; the stepping keeps it): the y pass of u, omega
  // skips the kz > kc lines and reads only the entries inside the mask radius
  // as well (the others are zero after the x pass)
  bool band = false;
  int* bandflag = nullptr;
  hipEvent_t evA = nullptr, evB = nullptr;
  C<T>*uk = nullptr, *us = nullptr, *acc = nullptr, *w6 = nullptr, *wn = nullptr;
  T* ubuf[2] = {nullptr, nullptr};       // planar physical u (API / plain paths)
  // droplet gather records {u, 0} and {v, n_stage} per node (z-quad layout,
  // recU / recV), one array per stage parity: the density chain of stage s
  // reads grec[s & 1] while the flow chain of stage s + 1 writes the other
  V4<T>* grec[2] = {nullptr, nullptr};
  T *v = nullptr, *n_state = nullptr, *n_stage = nullptr, *nacc = nullptr, *fc = nullptr;
  double *partials = nullptr, *partials2 = nullptr, *clip_stage = nullptr, *clip_end = nullptr, *clip_none = nullptr;
  double *maxpart = nullptr, *ek = nullptr, *fac = nullptr, *lastmax = nullptr;
  int *flag = nullptr, *shells = nullptr;
  int smax = 0;
  int64_t cap = 0, np = 0;
  // Ensemble of B realizations: every buffer allocated through arena() has
  // one copy per realization, rs bytes apart (a per-realization slab), and the
  // step's launches put the realization in blockIdx.z (rsel() in the kernels).
  int B = 1;
  size_t rs = 0;
  struct ArenaSlot { void** ptr; size_t off; };
  std::vector<ArenaSlot> arena_plan;
  size_t arena_bytes = 0;
  double *x0 = nullptr, *sx = nullptr;
  PT *kxp = nullptr;                 // stage dx/dt = stage velocity, held exactly in PT
  PT *v0 = nullptr, *sv = nullptr, *kvp = nullptr;
  double *rows_a = nullptr, *rows_b = nullptr, *wrow = nullptr;
  // spatial sort: slot -> input row, bin keys/ranks, bin starts, double buffers
  int64_t *rid = nullptr, *rid2 = nullptr;
  int *pkey = nullptr, *prank = nullptr, *hist = nullptr, *bsum = nullptr;
  double* x0b = nullptr;
  PT* v0b = nullptr;
  double4* dep = nullptr;            // per-slot stage position and W for the brick deposit
  int *flist = nullptr, *fcount = nullptr;
  unsigned long long* fx = nullptr;  // fixed-point scratch field of the fallback deposit (zero between steps)
  unsigned long long* wmaxp = nullptr;   // largest W of a droplet stage (fixed-point deposit bound)
  bool sort_enabled = true;
  // y and z passes of a chain run slab by slab (x-planes) so the y<->z
  // intermediate of one slab stays in the Infinity Cache; 1 = whole field.
  int nslab = 1;
  bool sorted_ok = false;            // bin starts (hist) describe the current slots
  static constexpr int NB = Bins<N>::NB;
  static constexpr int NBS = (NB + 4095) / 4096;
  DropPhys dp;
  // kernel timing (ld_profile): event pairs around launches of one kernel class
  int prof_id = -1;
  std::vector<hipEvent_t> prof_ev;
  size_t prof_used = 0;
  static constexpr size_t PROF_MAX = 8192;

  explicit Engine(const ld_params& prm) : p(prm) {}
  ~Engine() override {
    if (st2) { (void)hipStreamSynchronize(st2); (void)hipStreamDestroy(st2); }
    if (stg) { (void)hipStreamSynchronize(stg); (void)hipStreamDestroy(stg); }
    if (stf) { (void)hipStreamSynchronize(stf); (void)hipStreamDestroy(stf); }
    for (StepGraph& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    if (evA) (void)hipEventDestroy(evA);
    if (evG0) (void)hipEventDestroy(evG0);
    if (evG1) (void)hipEventDestroy(evG1);
    if (evB) (void)hipEventDestroy(evB);
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
  // Reserve a per-realization buffer in the slab; arena_commit() allocates
  // B slabs and sets every reserved pointer to realization 0's copy.
  template <class X> void arena(X*& ptr, size_t count) {
    if (count == 0) count = 1;
    arena_bytes = (arena_bytes + 255) / 256 * 256;
    arena_plan.push_back({reinterpret_cast<void**>(&ptr), arena_bytes});
    arena_bytes += count * sizeof(X);
  }
  int arena_commit() {
    rs = (arena_bytes + 255) / 256 * 256;
    void* base = nullptr;
    LD_HIP(hipMalloc(&base, rs * (size_t)B));
    allocs.push_back(base);
    for (const ArenaSlot& a : arena_plan) *a.ptr = static_cast<char*>(base) + a.off;
    return LD_OK;
  }
  // realization b's copy of an arena buffer (host-side address arithmetic)
  template <class X> X* at(X* ptr, int b) const {
    return reinterpret_cast<X*>(reinterpret_cast<char*>(ptr) + (size_t)b * rs);
  }
  int batch() const override { return B; }

  // Autotuner: time the 6-field y-inverse pass with M, M/2 and 2M lines per
  // block, then each of the z passes of u x omega and v with its default and
  // half tile height and each of the three x passes with its default and half
  // lines per block (on the context's zeroed work buffers, as the truncated
  // step launches them; after one warm-up launch each, the median of five)
  // and keep the fastest of each.
  double tune_us[3] = {0.0, 0.0, 0.0};
  template <class F> int time_us(hipEvent_t e0, hipEvent_t e1, F&& launch, double& us) {
    std::vector<float> t;
    for (int r = 0; r < 6; ++r) {
      LD_HIP(hipEventRecord(e0, st));
      LD_TRY(launch());
      LD_HIP(hipEventRecord(e1, st));
      LD_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      LD_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (r > 0) t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    us = 1000.0 * t[t.size() / 2];
    return LD_OK;
  }
  int autotune_geometry(hipEvent_t e0, hipEvent_t e1) {
    LD_TRY(fill<C<T>>(st, wn, 4 * CF, C<T>{}));
    LD_TRY(fill<C<T>>(st, us, 3 * CF, C<T>{}));
    LD_TRY(fill<C<T>>(st, uk, 3 * CF, C<T>{}));
    LD_TRY(fill<C<T>>(st, acc, 3 * CF, C<T>{}));
    LD_TRY(fill<T>(st, n_state, RF, (T)0));
    const bool alt[5] = {P6H != P6, P3H != P3, M3H != M3, M3H != M3, M4H != M4};
    const bool dbl[5] = {P6D != P6, P3D != P3, false, false, false};
    RhsArgs<T> ra;
    ra.nl = w6; ra.us = us; ra.ubase = uk; ra.acc = acc; ra.usnext = us; ra.ku = nullptr;
    ra.fac = fac; ra.fs = CF; ra.nu_eff = (T)(p.nu * p.isexplicit); ra.tps = (T)p.tps;
    ra.wacc = (T)0; ra.anext = (T)0; ra.scale = inv(); ra.acc_init = 0; ra.kcut = kc; ra.band = kc < N / 2 ? 1 : 0;
    for (int k = 0; k < 5; ++k) {
      for (int v = 0; v < 3; ++v) {
        vus[k][v] = 0.0;
        if ((v == 1 && !alt[k]) || (v == 2 && !dbl[k])) continue;
        vsel[k] = v;
        auto launch = [&]() -> int {
          switch (k) {
            case 0: return zcross(st, nullptr, grec[0], 0, N, kc);
            case 1: return zvel(st, nullptr, grec[0], n_state, clip_none, 0, N);
            case 2: return nonlinear_x(st, uk, kc);
            case 3: return xrhs(st, ra);
            default: return xkn(st);
          }
        };
        LD_TRY(time_us(e0, e1, launch, vus[k][v]));
      }
      int best = 0;
      for (int v = 1; v < 3; ++v)
        if (vus[k][v] > 0.0 && vus[k][v] < vus[k][best]) best = v;
      vsel[k] = best;
    }
    return LD_OK;
  }
  int autotune() {
    LD_TRY(fill<C<T>>(st, w6, 6 * CF, C<T>{}));
    hipEvent_t e0, e1;
    LD_HIP(hipEventCreate(&e0));
    LD_HIP(hipEventCreate(&e1));
    const int cand[3] = {M, MH, M2};
    const bool wide_ok = (N / nslab) % M2 == 0 && M2 * N1 <= 1024;
    for (int c = 0; c < 3; ++c) {
      if (c == 2 && !wide_ok) { tune_us[2] = 0.0; continue; }
      my = cand[c];
      std::vector<float> t;
      for (int r = 0; r < 6; ++r) {
        LD_HIP(hipEventRecord(e0, st));
        LD_TRY((pass<1, +1>(st, w6, w6, 6, (T)1)));
        LD_HIP(hipEventRecord(e1, st));
        LD_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LD_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) t.push_back(ms);
      }
      std::sort(t.begin(), t.end());
      tune_us[c] = 1000.0 * t[t.size() / 2];
    }
    int best = 0;
    for (int c = 1; c < 3; ++c)
      if (tune_us[c] > 0.0 && tune_us[c] < tune_us[best]) best = c;
    my = cand[best];
    // the other line split at the chosen lines, timed for the record only
    {
      const int keep = ysplit;
      ysplit = 1;
      auto launch = [&]() -> int { return pass<1, +1>(st, w6, w6, 6, (T)1); };
      LD_TRY(time_us(e0, e1, launch, split_us));
      ysplit = keep;
    }
    const int rc = autotune_geometry(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  int tuning_all(int32_t* choice, double* us_out) override {
    const int32_t sizes[5][3] = {{P6, P6H, P6D}, {P3, P3H, P3D}, {M3, M3H, M3}, {M3, M3H, M3}, {M4, M4H, M4}};
    if (choice) {
      choice[0] = my;
      for (int k = 0; k < 5; ++k) choice[1 + k] = sizes[k][vsel[k]];
      choice[6] = ysplit ? NQ : N1;
    }
    if (us_out) {
      for (int c = 0; c < 3; ++c) us_out[c] = tune_us[c];
      for (int k = 0; k < 5; ++k)
        for (int v = 0; v < 3; ++v) us_out[3 + 3 * k + v] = vus[k][v];
      us_out[18] = split_us;
    }
    return LD_OK;
  }
  int tuning(int32_t* ylines, double* us) override {
    if (ylines) *ylines = my;
    if (us) { us[0] = tune_us[0]; us[1] = tune_us[1]; us[2] = tune_us[2]; }
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
    if (p.batch < 0 || p.batch > 4096) { set_error("batch must be in [0, 4096]"); return LD_ERR_ARG; }
    B = p.batch > 1 ? p.batch : 1;
    // LD_SERIAL=1: the flow stream waits for each stage's density chain, so the
    // two chains never run concurrently (per-kernel timing without overlap)
    if (const char* e = getenv("LD_SERIAL")) serial = atoi(e) != 0;
    if (const char* e = getenv("LD_YSPLIT")) ysplit = atoi(e) != 0;
    if (const char* e = getenv("LD_LATE_DENSITY")) late_density = atoi(e) != 0;
    if (const char* e = getenv("LD_TRUNCATE")) kc = atoi(e) != 0 ? N / 3 : N / 2;
    if (const char* e = getenv("LD_PSTAGE_LDS")) {
      const int v = atoi(e);
      if (v >= 0 && v <= 65536) pstage_lds = (unsigned)v;
    }
    if (const char* e = getenv("LD_SLABS")) {
      const int v = atoi(e);
      // a slab must hold whole y-pass line groups and whole z tiles (all powers of two)
      const int zmax = P1 > P3D ? (P1 > P6D ? P1 : P6D) : (P3D > P6D ? P3D : P6D);
      if (v >= 1 && N % v == 0 && (N / v) % M == 0 && ((N / v) * N) % zmax == 0) nslab = v;
    }
    // twiddles exp(-2 pi i m / N), rounded from long double
    std::vector<C<T>> h(N);
    for (int m = 0; m < N; ++m) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)N;
      h[m] = mkc<T>((T)cosl(ang), (T)sinl(ang));
    }
    LD_TRY(alloc(tw, N));
    LD_HIP(hipMemcpy(tw, h.data(), N * sizeof(C<T>), hipMemcpyHostToDevice));
    // LD_CU_MASK=<8 hex words>: the density/droplet stream runs on the CUs of the
    // mask and the step's flow stream (internal) on the others
    std::vector<uint32_t> dmask;
    if (const char* e = getenv("LD_CU_MASK")) {
      const char* q = e;
      while (*q && dmask.size() < 16) {
        char* end = nullptr;
        dmask.push_back((uint32_t)strtoul(q, &end, 16));
        if (end == q) break;
        q = (*end == ',') ? end + 1 : end;
      }
    }
    if (!dmask.empty()) {
      std::vector<uint32_t> fmask(dmask.size());
      for (size_t i = 0; i < dmask.size(); ++i) fmask[i] = ~dmask[i];
      LD_HIP(hipExtStreamCreateWithCUMask(&st2, (uint32_t)dmask.size() * 32, dmask.data()));
      LD_HIP(hipExtStreamCreateWithCUMask(&stf, (uint32_t)fmask.size() * 32, fmask.data()));
    } else {
      LD_HIP(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));   // default priority
    }
    use_graph = p.use_graph != 0;
    if (use_graph) LD_HIP(hipStreamCreateWithFlags(&stg, hipStreamNonBlocking));
    if (use_graph || stf) {
      LD_HIP(hipEventCreateWithFlags(&evG0, hipEventDisableTiming));
      LD_HIP(hipEventCreateWithFlags(&evG1, hipEventDisableTiming));
    }
    LD_HIP(hipEventCreateWithFlags(&evA, hipEventDisableTiming));
    LD_HIP(hipEventCreateWithFlags(&evB, hipEventDisableTiming));
    // per-realization state and work buffers (one slab per realization)
    arena(uk, 3 * CF); arena(us, 3 * CF); arena(acc, 3 * CF); arena(w6, 6 * CF); arena(wn, 4 * CF);
    arena(ubuf[0], 3 * RF); arena(ubuf[1], 3 * RF); arena(v, 3 * RF);
    arena(grec[0], 2 * RF); arena(grec[1], 2 * RF);
    arena(n_state, RF); arena(n_stage, RF); arena(nacc, RF); arena(fc, RF);
    const int npart = NBLK_Z1 > NBLK_SUM ? NBLK_Z1 : NBLK_SUM;
    arena(partials, 2 * (size_t)npart); arena(partials2, 2 * (size_t)npart);
    arena(clip_stage, 4); arena(clip_end, 4); arena(clip_none, 4);
    arena(maxpart, NBLK_X); arena(ek, Nf); arena(fac, Nf); arena(lastmax, 1); arena(flag, 1);
    if (cap > 0) {
      arena(x0, 3 * cap); arena(sx, 3 * cap); arena(kxp, 3 * cap);
      arena(v0, 4 * cap); arena(sv, 4 * cap); arena(kvp, 4 * cap);
      arena(rows_a, 7 * cap); arena(rows_b, 7 * cap); arena(wrow, cap);
      arena(rid, cap); arena(rid2, cap); arena(pkey, cap); arena(prank, cap);
      arena(hist, NB); arena(bsum, NBS);
      arena(x0b, 3 * cap); arena(v0b, 4 * cap);
      arena(dep, cap); arena(flist, cap); arena(fcount, 1); arena(wmaxp, 1); arena(fx, RF);
    }
    LD_TRY(arena_commit());
    LD_TRY(alloc(shells, 8));
    LD_TRY(alloc(bandflag, 1));
    LD_HIP(hipMemcpy(shells, p.shell_no, 8 * sizeof(int), hipMemcpyHostToDevice));
    const double cn[4] = {1.0, 1.0, 0.0, 0.0};
    for (int b = 0; b < B; ++b) {
      LD_HIP(hipMemset(at(fac, b), 0, Nf * sizeof(double)));
      LD_HIP(hipMemset(at(ek, b), 0, Nf * sizeof(double)));
      LD_HIP(hipMemset(at(flag, b), 0, sizeof(int)));
      LD_HIP(hipMemset(at(lastmax, b), 0, sizeof(double)));
      LD_HIP(hipMemcpy(at(clip_none, b), cn, sizeof(cn), hipMemcpyHostToDevice));
      LD_HIP(hipMemcpy(at(clip_stage, b), cn, sizeof(cn), hipMemcpyHostToDevice));
      LD_HIP(hipMemcpy(at(clip_end, b), cn, sizeof(cn), hipMemcpyHostToDevice));
    }
    if (cap > 0) {
      LD_TRY(fill<unsigned long long>(st2, fx, RF, 0ull));
      LD_HIP(hipStreamSynchronize(st2));
    }
    static_assert(NBS <= 4096, "bin-start scan supports at most 4096 partial sums");
    dp.L = p.L;
    dp.dx = p.L / N;                       // np.linspace(0, L, N, endpoint=False) step
    dp.pi_2dx = M_PI / (2.0 * dp.dx);
    dp.inv_vol = 1.0 / (dp.dx * dp.dx * dp.dx);
    dp.factor = p.factor; dp.growth = p.growthfactor; dp.st_s = p.st_s; dp.tau_eta = p.tau_eta; dp.g = p.g;
    dp.invL = 1.0 / p.L;
    dp.inv_dx = (double)N / p.L;
    dp.inv_factor = 1.0 / p.factor;
    if (p.autotune) LD_TRY(autotune());
    // LD_TUNE_FORCE=a,b,c,d,e (0/1 each): force the geometry choices (measurement, tests)
    if (const char* e = getenv("LD_TUNE_FORCE")) {
      for (int k = 0; k < 5 && *e; ++k) {
        vsel[k] = (*e == '1') ? 1 : ((*e == '2' && k < 2) ? 2 : 0);
        while (*e && *e != ',') ++e;
        if (*e == ',') ++e;
      }
    }
    return LD_OK;
  }

  void set_stream(hipStream_t s) override { st = s; }

  // ------------------------------------------------------------ launch helpers
  // Line pass over x-planes [x0, x0 + nx) (AXIS 1) or all lines (AXIS 0).
  // y passes run with my lines per block: M, M/2 or 2M (2M needs N/nslab to be
  // a multiple of 2M), the autotuner's choice; the lines a block owns do not
  // change the arithmetic.
  static constexpr int MH = M / 2, M2 = 2 * M;
  int my = M;
  // ysplit = 1 (LD_YSPLIT=1): y passes split each line N = (N1/2) (2 N2)
  // instead of N1 N2.  Another factorisation rounds differently, so this is
  // a measurement switch, not an autotuner choice (the autotuner only chooses
  // between launch shapes that give identical bits); the autotuner times it.
  static constexpr int NQ = N1 / 2;
  int ysplit = 0;
  double split_us = 0.0;
  template <int AXIS, int DIR> int pass(hipStream_t s, const C<T>* in, C<T>* out, int nf, T scale, int x0 = 0,
                                        int nx = N, int kcut = N / 2, int r2 = -1) {
    const int kid = AXIS == 1 ? LD_K_YPASS : -2;
    // y passes: only the lines with kz <= kcut, lw of them per x-plane
    const int lw = (kcut < N / 2 ? kcut : N / 2) + 1;
    const int qoff = AXIS == 1 ? x0 * lw : 0, qend = AXIS == 1 ? (x0 + nx) * lw : 0;
    const int nl = nx * lw;
    LD_TRY(prof_mark(kid, s));
    if (AXIS == 1 && ysplit) {
      if (my == MH)
        hipLaunchKernelGGL((k_pass<T, N, MH, AXIS, DIR, NQ>), dim3((nl + MH - 1) / MH, nf, B), dim3(MH * NQ), 0, s,
                           in, out, CF, (const C<T>*)tw, scale, qoff, rs, lw, qend, r2);
      else if (my == M2)
        hipLaunchKernelGGL((k_pass<T, N, M2, AXIS, DIR, NQ>), dim3((nl + M2 - 1) / M2, nf, B), dim3(M2 * NQ), 0, s,
                           in, out, CF, (const C<T>*)tw, scale, qoff, rs, lw, qend, r2);
      else
        hipLaunchKernelGGL((k_pass<T, N, M, AXIS, DIR, NQ>), dim3((nl + M - 1) / M, nf, B), dim3(M * NQ), 0, s,
                           in, out, CF, (const C<T>*)tw, scale, qoff, rs, lw, qend, r2);
    } else if (AXIS == 1 && my == MH)
      hipLaunchKernelGGL((k_pass<T, N, MH, AXIS, DIR>), dim3((nl + MH - 1) / MH, nf, B), dim3(MH * N1), 0, s, in, out,
                         CF, (const C<T>*)tw, scale, qoff, rs, lw, qend, r2);
    else if (AXIS == 1 && my == M2)
      hipLaunchKernelGGL((k_pass<T, N, M2, AXIS, DIR>), dim3((nl + M2 - 1) / M2, nf, B), dim3(M2 * N1), 0, s, in, out,
                         CF, (const C<T>*)tw, scale, qoff, rs, lw, qend, r2);
    else
      hipLaunchKernelGGL((k_pass<T, N, M, AXIS, DIR>), dim3(AXIS == 1 ? (nl + M - 1) / M : NBLK_X, nf, B), dim3(M * N1),
                         0, s, in, out, CF, (const C<T>*)tw, scale, qoff, rs, lw, qend, r2);
    LD_TRY(prof_mark(kid, s));
    LD_LAUNCHED();
    return LD_OK;
  }
  // z pass over the (x, y) lines of x-planes [x0, x0 + nx); input entries
  // with kz > kcin are zero (not loaded), output entries with kz > kcout are
  // not needed.
  template <int P, int FIN, int FOUT, class Op> int zrun(hipStream_t s, const Op& op, int kid = -2, int x0 = 0,
                                                         int nx = N, int kcout = N / 2, int kcin = N / 2) {
    LD_TRY(prof_mark(kid, s));
    hipLaunchKernelGGL((k_z<T, N, P, FIN, FOUT, Op>), dim3(nx * N / P, 1, B), dim3(ZThreads<N, P, FIN, FOUT>::v), 0,
                       s, op, (const C<T>*)tw, x0 * N, rs, kcin, kcout);
    LD_TRY(prof_mark(kid, s));
    LD_LAUNCHED();
    return LD_OK;
  }
  template <class X> int fill(hipStream_t s, X* x, size_t n, X val) {
    // grid-stride fill; small fills (counters, flags) take as few blocks as they need
    const size_t nb = (n + 255) / 256;
    hipLaunchKernelGGL((k_fill<X>), dim3(nb < 1024 ? (unsigned)(nb ? nb : 1) : 1024u, 1, B), dim3(256), 0, s, x, n,
                       val, rs);
    LD_LAUNCHED();
    return LD_OK;
  }
  int shell(hipStream_t s, const C<T>* uin, double fscale) {
    const double norm0 = 1.0 / ((double)N * N * N * N * N * N / std::pow(2.0 * M_PI, 3.0));
    if (p.nshells == 0) return LD_OK;
    hipLaunchKernelGGL((k_shell<T, N>), dim3(1, 1, B), dim3(256), 0, s, uin, CF, smax, (const int*)shells, p.nshells,
                       p.f0, fscale, norm0, ek, fac, rs);
    LD_LAUNCHED();
    return LD_OK;
  }
  int clip_sums(hipStream_t s, const T* x, double* clip) {
    hipLaunchKernelGGL((k_sum2<T>), dim3(NBLK_SUM, 1, B), dim3(256), 0, s, x, RF, partials, rs);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_clip_final, dim3(1, 1, B), dim3(256), 0, s, (const double*)partials, NBLK_SUM,
                       1.0 / ((double)N * N * N), clip, rs);
    LD_LAUNCHED();
    return LD_OK;
  }
  // Event pair around one launch of kernel class `id` (when it is profiled).
  int prof_mark(int id, hipStream_t s) {
    if (id != prof_id) return LD_OK;
    if (prof_used >= prof_ev.size()) {
      if (prof_ev.size() >= PROF_MAX) return LD_OK;   // pool exhausted: stop recording
      hipEvent_t e;
      LD_HIP(hipEventCreate(&e));
      prof_ev.push_back(e);
    }
    LD_HIP(hipEventRecord(prof_ev[prof_used++], s));
    return LD_OK;
  }
  int profile(int kernel, int enable) override {
    prof_id = enable ? kernel : -1;
    prof_used = 0;
    return LD_OK;
  }
  int profile_read(double* total_ms, int64_t* launches) override {
    LD_HIP(hipDeviceSynchronize());
    double tot = 0.0;
    const size_t pairs = prof_used / 2;
    for (size_t i = 0; i < pairs; ++i) {
      float ms = 0.f;
      LD_HIP(hipEventElapsedTime(&ms, prof_ev[2 * i], prof_ev[2 * i + 1]));
      tot += ms;
    }
    *total_ms = tot;
    *launches = (int64_t)pairs;
    prof_used = 0;
    return LD_OK;
  }

  // make `waiter` wait for all work so far on `signaller`
  int join(hipStream_t waiter, hipStream_t signaller, hipEvent_t ev) {
    if (waiter == signaller) return LD_OK;
    LD_HIP(hipEventRecord(ev, signaller));
    LD_HIP(hipStreamWaitEvent(waiter, ev, 0));
    return LD_OK;
  }
  static T inv() { return (T)(1.0 / N); }

  // ------------------------------------------------------------ transforms
  int rfft(const void* uin, void* fuout, int nf) override {
    const T* src = (const T*)uin;
    C<T>* dst = (C<T>*)fuout;
    for (int f = 0; f < nf; ++f) {
      OpZFwd<T, N, P1, 1> op{src + f * RF, dst + f * CF, RF, CF};
      LD_TRY((zrun<P1, 0, 1>(st, op)));
    }
    LD_TRY((pass<1, -1>(st, dst, dst, nf, (T)1)));
    LD_TRY((pass<0, -1>(st, dst, dst, nf, (T)1)));
    return LD_OK;
  }
  int irfft(const void* fuin, void* uout, int nf) override {
    const C<T>* src = (const C<T>*)fuin;
    T* dst = (T*)uout;
    for (int f0 = 0; f0 < nf; f0 += 6) {
      const int k = nf - f0 < 6 ? nf - f0 : 6;
      LD_TRY((pass<0, +1>(st, src + f0 * CF, w6, k, inv())));
      LD_TRY((pass<1, +1>(st, w6, w6, k, inv())));
      for (int f = 0; f < k; ++f) {
        OpZInv<T, N, P1, 1> op{w6 + f * CF, dst + (f0 + f) * RF, CF, RF, inv()};
        LD_TRY((zrun<P1, 1, 0>(st, op)));
      }
    }
    return LD_OK;
  }

  int forcing(const void* ukin, void* fkout, double* ek_out) override {
    LD_TRY(shell(st, (const C<T>*)ukin, p.isforcing));
    const size_t nb = (CF + 255) / 256;
    hipLaunchKernelGGL((k_forcing_apply<T, N>), dim3(nb), dim3(256), 0, st, (const C<T>*)ukin, (C<T>*)fkout, CF,
                       (const double*)fac);
    LD_LAUNCHED();
    if (ek_out) LD_HIP(hipMemcpyAsync(ek_out, ek, Nf * sizeof(double), hipMemcpyDeviceToDevice, st));
    return LD_OK;
  }

  int clip_zero(void* x, void* out) override {
    LD_TRY(clip_sums(st, (const T*)x, clip_end));
    hipLaunchKernelGGL((k_clip_apply<T>), dim3(1024), dim3(256), 0, st, (T*)x, (T*)out, RF, (const double*)clip_end, 1,
                       (size_t)0);
    LD_LAUNCHED();
    return LD_OK;
  }

  // ------------------------------------------------------------ stage pieces
  // Flow chain on stream s: u, omega -> physical u (uphys), z/y-forward of
  // u x omega in w6[0..3).
  int nonlinear(hipStream_t s, const C<T>* usp, T* uphys, V4<T>* g = nullptr) {
    LD_TRY(nonlinear_x(s, usp));
    return nonlinear_yz(s, uphys, g);
  }
  // kcu < N/2: the stage u_hat has no nonzero mode outside the mask
  int nonlinear_x(hipStream_t s, const C<T>* usp, int kcu = N / 2) {
    LD_TRY(prof_mark(LD_K_XUW, s));
    if (vsel[2])
      hipLaunchKernelGGL((k_x_uw3<T, N, M3H>), dim3(N * Nf / M3H, 1, B), dim3(3 * M3H * N1), 0, s, usp, w6, CF,
                         (const C<T>*)tw, inv(), rs, kcu);
    else
      hipLaunchKernelGGL((k_x_uw3<T, N, M3>), dim3(NBLK_X3, 1, B), dim3(3 * M3 * N1), 0, s, usp, w6, CF,
                         (const C<T>*)tw, inv(), rs, kcu);
    LD_LAUNCHED();
    LD_TRY(prof_mark(LD_K_XUW, s));
    return LD_OK;
  }
  // kcu < N/2: the stage u_hat has no nonzero mode outside the mask, so the
  // y-inverse output (the input here) is zero at kz > kcu
  int zcross(hipStream_t s, T* uphys, V4<T>* g, int x0, int nx, int kcu) {
    if (vsel[0] == 2) {
      OpZCross<T, N, P6D> oc{w6, CF, uphys, g, inv()};
      return zrun<P6D, 6, 3>(s, oc, LD_K_ZCROSS, x0, nx, kc, kcu);
    }
    if (vsel[0]) {
      OpZCross<T, N, P6H> oc{w6, CF, uphys, g, inv()};
      return zrun<P6H, 6, 3>(s, oc, LD_K_ZCROSS, x0, nx, kc, kcu);
    }
    OpZCross<T, N, P6> oc{w6, CF, uphys, g, inv()};
    return zrun<P6, 6, 3>(s, oc, LD_K_ZCROSS, x0, nx, kc, kcu);
  }
  int nonlinear_yz(hipStream_t s, T* uphys, V4<T>* g, int kcu = N / 2) {
    const int nx = N / nslab;
    for (int x0 = 0; x0 < N; x0 += nx) {
      LD_TRY((pass<1, +1>(s, w6, w6, 6, inv(), x0, nx, kcu, kcu < N / 2 ? sphere() : -1)));
      LD_TRY(zcross(s, uphys, g, x0, nx, kcu));
      LD_TRY((pass<1, -1>(s, w6, w6, 3, (T)1, x0, nx, kc, sphere())));
    }
    return LD_OK;
  }
  // v_hat (x-inverse in w6[0..3)): y-inverse; then z-pass to physical v, and n v
  // z-forward into wn[0..3).
  // v_hat y-inverse and z pass (v out, n v forward) slab by slab; `wait`
  // (may be null) is waited for before the first z pass writes v / nv.
  int zvel(hipStream_t s, T* vphys, V4<T>* g, const T* ns, const double* clip, int x0, int nx) {
    if (vsel[1] == 2) {
      OpZVel<T, N, P3D> ov{w6, wn, CF, vphys, g, ns, clip, inv(), (T)(p.tps * p.g)};
      return zrun<P3D, 3, 3>(s, ov, LD_K_ZVEL, x0, nx, kc, kc);
    }
    if (vsel[1]) {
      OpZVel<T, N, P3H> ov{w6, wn, CF, vphys, g, ns, clip, inv(), (T)(p.tps * p.g)};
      return zrun<P3H, 3, 3>(s, ov, LD_K_ZVEL, x0, nx, kc, kc);
    }
    OpZVel<T, N, P3> ov{w6, wn, CF, vphys, g, ns, clip, inv(), (T)(p.tps * p.g)};
    return zrun<P3, 3, 3>(s, ov, LD_K_ZVEL, x0, nx, kc, kc);
  }
  int velocity(hipStream_t s, T* vphys, V4<T>* g, const T* ns, const double* clip, hipEvent_t wait) {
    const int nx = N / nslab;
    for (int x0 = 0; x0 < N; x0 += nx) {
      LD_TRY((pass<1, +1>(s, w6, w6, 3, inv(), x0, nx, kc, sphere())));
      if (x0 == 0 && wait) LD_HIP(hipStreamWaitEvent(s, wait, 0));
      LD_TRY(zvel(s, vphys, g, ns, clip, x0, nx));
    }
    return LD_OK;
  }
  int xkn(hipStream_t s) {
    if (vsel[4])
      hipLaunchKernelGGL((k_x_kn4<T, N, M4H>), dim3(N * Nf / M4H, 1, B), dim3(4 * M4H * N1), 0, s, wn, CF,
                         (const C<T>*)tw, inv(), rs, kc);
    else
      hipLaunchKernelGGL((k_x_kn4<T, N, M4>), dim3(NBLK_X4, 1, B), dim3(4 * M4 * N1), 0, s, wn, CF, (const C<T>*)tw,
                         inv(), rs, kc);
    LD_LAUNCHED();
    return LD_OK;
  }
  int xrhs(hipStream_t s, const RhsArgs<T>& ra) {
    if (vsel[3])
      hipLaunchKernelGGL((k_x_rhs3<T, N, M3H>), dim3(N * Nf / M3H, 1, B), dim3(3 * M3H * N1), 0, s, ra,
                         (const C<T>*)tw, rs);
    else
      hipLaunchKernelGGL((k_x_rhs3<T, N, M3>), dim3(NBLK_X3, 1, B), dim3(3 * M3 * N1), 0, s, ra, (const C<T>*)tw, rs);
    LD_LAUNCHED();
    return LD_OK;
  }
  // Density chain on stream s: fc -> z-forward into wn[3]; y-forward of
  // wn[0..4); kn_hat x-inverse in wn[0]; y-inverse.
  // (leaves kn_hat y-inverse-transformed in wn[0] when kz == nullptr; with
  // kz the z pass runs per slab too)
  template <class OpKn> int density_rhs(hipStream_t s, const T* fcin, const OpKn* kz) {
    OpZFwd<T, N, P1, 1> of{fcin, wn + 3 * CF, RF, CF};
    const int nx = N / nslab;
    for (int x0 = 0; x0 < N; x0 += nx) {
      LD_TRY((zrun<P1, 0, 1>(s, of, -2, x0, nx, kc)));
      LD_TRY((pass<1, -1>(s, wn, wn, 4, (T)1, x0, nx, kc, sphere())));
    }
    LD_TRY(prof_mark(LD_K_XKN, s));
    LD_TRY(xkn(s));
    LD_TRY(prof_mark(LD_K_XKN, s));
    for (int x0 = 0; x0 < N; x0 += nx) {
      LD_TRY((pass<1, +1>(s, wn, wn, 1, inv(), x0, nx, kc, sphere())));
      if (kz) LD_TRY((zrun<P1, 1, 0>(s, *kz, LD_K_ZKN, x0, nx, N / 2, kc)));
    }
    return LD_OK;
  }

  // One RK4 stage s (0..3) of the fused step on the context state.
  //   stream st  (flow):    forcing factors, u/omega, nonlinear term, RK update
  //                         of u_hat, v_hat, v (waits for the previous
  //                         stage's density chain before writing v / nv);
  //   stream st2 (density): droplets (gather, RHS, deposit, RK update),
  //                         fc and n v transforms, kn, RK update of n.
  // The next stage's flow chain depends only on u_hat, so it overlaps this
  // stage's density chain.  u (physical) is double-buffered for that.
  // One RK stage in three launch groups: the flow stage's x pass (a), the rest
  // of the flow stage (b) and the density/droplet chain (d).  The density chain
  // of stage s needs the flow stage s done and must finish before the flow
  // stage s + 1's velocity pass.  With late_density (the default for N >= 128)
  // the chain of stage s is queued after flow stage s + 1's x pass (stage_a),
  // so that x pass, whose blocks fill a CU each, does not compete with the
  // droplet kernels; the chain then overlaps the y and z passes of flow stage
  // s + 1.  Otherwise it is queued right after flow stage s.
  static constexpr double a_coef[5] = {0.0, 0.5, 0.5, 1.0, 0.0};
  int stage_a(int s) {
    const C<T>* usp = (s == 0) ? uk : us;
    LD_TRY(shell(st, usp, p.isforcing));
    return nonlinear_x(st, usp, band ? kc : N / 2);
  }
  int stage_b(int s, double h) {
    const C<T>* usp = (s == 0) ? uk : us;
    const T* ns = (s == 0) ? n_state : n_stage;
    V4<T>* gup = grec[s & 1];
    const double wh = (s == 0 || s == 3) ? h / 6.0 : h / 3.0;   // forced-dns-sm.py:342-358
    const double anext = a_coef[s + 1] * h;
    LD_TRY(nonlinear_yz(st, nullptr, gup, band ? kc : N / 2));
    RhsArgs<T> ra;
    ra.nl = w6; ra.us = usp; ra.ubase = uk; ra.acc = acc; ra.usnext = (s < 3) ? us : nullptr; ra.ku = nullptr;
    ra.fac = fac; ra.fs = CF; ra.nu_eff = (T)(p.nu * p.isexplicit); ra.tps = (T)p.tps;
    ra.wacc = (T)wh; ra.anext = (T)anext; ra.scale = inv(); ra.acc_init = (s == 0);
    ra.kcut = kc; ra.band = band ? 1 : 0;
    LD_TRY(prof_mark(LD_K_XRHS, st));
    LD_TRY(xrhs(st, ra));
    LD_TRY(prof_mark(LD_K_XRHS, st));
    // v, nv need the previous density chain done (n_s, clip, wn and v free)
    return velocity(st, nullptr, gup, ns, clip_stage, evB);
  }
  int stage_d(int s, double h) {
    V4<T>* gup = grec[s & 1];
    const double wh = (s == 0 || s == 3) ? h / 6.0 : h / 3.0;
    const double anext = a_coef[s + 1] * h;
    LD_TRY(join(st2, st, evA));
    if (!sorted_ok || np == 0) LD_TRY(fill<T>(st2, fc, RF, (T)0));
    if (np > 0) {
      LD_TRY(fill<int>(st2, fcount, 1, 0));
      LD_TRY(fill<unsigned long long>(st2, wmaxp, 1, 0ull));
      PStage<T, PT> ps;
      ps.np = (int)np; ps.cap = (int)cap; ps.stage = s;
      ps.x0 = x0; ps.v0 = v0; ps.sx = sx; ps.sv = sv; ps.kx = kxp; ps.kv = kvp;
      ps.gu = gup; ps.gv = gup; ps.dep = dep; ps.flist = flist; ps.fcount = fcount;
      ps.all_fallback = sorted_ok ? 0 : 1;
      ps.wmax = wmaxp;
      ps.ah = a_coef[s] * h; ps.wh = wh; ps.P = dp; ps.rs = rs;
      // 4 lanes per droplet; block count a multiple of 8 for the XCD mapping
      const unsigned nb = (unsigned)(((4 * np + 255) / 256 + 7) / 8 * 8);
      LD_TRY(prof_mark(LD_K_DROPLETS, st2));
      // pstage_lds: dynamic LDS reserved per block, which caps the blocks per CU
      // (and so the droplets in flight per XCD, whose stencils share its L2)
      if (B == 1)
        hipLaunchKernelGGL((k_particles_stage<T, PT, N, false>), dim3(nb), dim3(256), pstage_lds, st2, ps);
      else
        hipLaunchKernelGGL((k_particles_stage<T, PT, N, true>), dim3(nb, 1, B), dim3(256), pstage_lds, st2, ps);
      LD_LAUNCHED();
      LD_TRY(prof_mark(LD_K_DROPLETS, st2));
      LD_TRY(prof_mark(LD_K_DEPOSIT, st2));
      if (sorted_ok) {
        hipLaunchKernelGGL((k_brick_deposit<T, N>), dim3(Brick<N>::NB * Brick<N>::NB * Brick<N>::NB, 1, B),
                           dim3(Brick<N>::NT), 0, st2, (const double4*)dep, (const int*)hist, (int)np,
                           (const unsigned long long*)wmaxp, fc, dp, rs);
        LD_LAUNCHED();
      }
      const unsigned fb = sorted_ok ? 64u : (unsigned)((np + 255) / 256);
      hipLaunchKernelGGL((k_fallback_accum<T, N>), dim3(fb, 1, B), dim3(256), 0, st2, (const double4*)dep,
                         (const int*)flist, (const int*)fcount, (const unsigned long long*)wmaxp, fx, dp, rs);
      LD_LAUNCHED();
      hipLaunchKernelGGL((k_fallback_fold<T, N>), dim3(fb, 1, B), dim3(256), 0, st2, (const double4*)dep,
                         (const int*)flist, (const int*)fcount, (const unsigned long long*)wmaxp, fx, fc, dp, rs);
      LD_LAUNCHED();
      LD_TRY(prof_mark(LD_K_DEPOSIT, st2));
    }
    OpZKn<T, N, P1> ok;
    ok.w = wn; ok.fs = CF; ok.scale = inv(); ok.kn_out = nullptr;
    ok.n0 = n_state; ok.nacc = nacc; ok.nnext = (s < 3) ? n_stage : nullptr;
    ok.wacc = (T)wh; ok.anext = (T)anext; ok.acc_init = (s == 0);
    ok.sum_target = (s < 3) ? 0 : 1; ok.partials = partials2;
    LD_TRY(density_rhs(st2, fc, &ok));
    hipLaunchKernelGGL(k_clip_final, dim3(1, 1, B), dim3(256), 0, st2, (const double*)partials2, NBLK_Z1,
                       1.0 / ((double)N * N * N), (s < 3) ? clip_stage : clip_end, rs);
    LD_LAUNCHED();
    LD_HIP(hipEventRecord(evB, st2));
    if (serial) LD_TRY(join(st, st2, evB));
    return LD_OK;
  }

  // Stage-1 clip_zero(n) sums (forced-dns-sm.py:341) on the density stream.
  int begin_rk4() {
    LD_TRY(clip_sums(st2, n_state, clip_stage));
    LD_HIP(hipEventRecord(evB, st2));
    if (serial) LD_TRY(join(st, st2, evB));
    return LD_OK;
  }

  // RK4 over the context state: acc <- uknew, nacc <- nnew, droplet base <- sump.
  int rk4_stages(double h) {
    LD_TRY(begin_rk4());
    for (int s = 0; s < 4; ++s) {
      LD_TRY(stage_a(s));
      if (late_density && s > 0) LD_TRY(stage_d(s - 1, h));
      LD_TRY(stage_b(s, h));
      if (!late_density) LD_TRY(stage_d(s, h));
    }
    if (late_density) LD_TRY(stage_d(3, h));
    return LD_OK;
  }

  // Step epilogue (forced-dns-sm.py:583-620): n = clip_zero(nnew) on the
  // density stream; reality round trip, projection and blow-up test on the
  // flow stream.
  int epilogue() {
    hipLaunchKernelGGL((k_clip_apply<T>), dim3(1024, 1, B), dim3(256), 0, st2, nacc, n_state, RF, (const double*)clip_end,
                       1, rs);
    LD_LAUNCHED();
    // reality round trip + projection + blow-up maxima in one spectral pass
    hipLaunchKernelGGL((k_epi_proj<T, N>), dim3(NBLK_X, 1, B), dim3(256), 0, st, (const C<T>*)acc, uk, CF, maxpart, rs,
                       band ? sphere() : -1);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_max_final, dim3(1, 1, B), dim3(256), 0, st, (const double*)maxpart, NBLK_X,
                       100.0 * (double)N * N * N, flag, lastmax, rs);
    LD_LAUNCHED();
    return LD_OK;
  }

  // ------------------------------------------------------------ state
  // Droplet rows of every realization: coord holds B blocks of n_ rows.
  int set_rows(const double* coord, int64_t n_) {
    sorted_ok = false;
    if (n_ > cap) { set_error("%lld droplets exceed the context capacity %lld", (long long)n_, (long long)cap); return LD_ERR_CAPACITY; }
    np = n_;
    if (np > 0) {
      for (int b = 0; b < B; ++b) {
        hipLaunchKernelGGL((k_rows_to_soa<PT>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                           coord + (size_t)b * np * 7, (int)np, (int)cap, at(x0, b), at(v0, b));
        LD_LAUNCHED();
        hipLaunchKernelGGL(k_iota64, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, at(rid, b), (int)np);
        LD_LAUNCHED();
      }
    }
    return LD_OK;
  }
  int get_rows(double* coord) {
    if (np > 0) {
      for (int b = 0; b < B; ++b) {
        hipLaunchKernelGGL((k_soa_to_rows<PT>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                           (const double*)at(x0, b), (const PT*)at(v0, b), (const int64_t*)at(rid, b), (int)np,
                           (int)cap, coord + (size_t)b * np * 7);
        LD_LAUNCHED();
      }
    }
    return LD_OK;
  }
  // uk, n and coord hold B realizations back to back.
  int state_set(const void* ukin, const void* nin, const double* coord, int64_t n_) override {
    LD_TRY(join(st, st2, evB));
    for (int b = 0; b < B; ++b) {
      LD_HIP(hipMemcpyAsync(at(uk, b), (const C<T>*)ukin + (size_t)b * 3 * CF, 3 * CF * sizeof(C<T>),
                            hipMemcpyDeviceToDevice, st));
      LD_HIP(hipMemcpyAsync(at(n_state, b), (const T*)nin + (size_t)b * RF, RF * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    LD_TRY(check_band());
    return set_rows(coord, n_);
  }
  // band <- no realization's u_hat has a nonzero mode outside the mask (one
  // synchronisation, when the state is set)
  int check_band() {
    band = false;
    if (kc >= N / 2) return LD_OK;
    hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(256), 0, st, bandflag, (size_t)1, 0, (size_t)0);
    LD_LAUNCHED();
    for (int b = 0; b < B; ++b) {
      hipLaunchKernelGGL((k_band_check<T, N>), dim3(1024), dim3(256), 0, st, (const C<T>*)at(uk, b),
                         (size_t)3 * CF, sphere(), bandflag);
      LD_LAUNCHED();
    }
    int h = 1;
    LD_HIP(hipMemcpyAsync(&h, bandflag, sizeof(int), hipMemcpyDeviceToHost, st));
    LD_HIP(hipStreamSynchronize(st));
    band = h == 0;
    // the RK update skips the lines outside the mask radius of such a state:
    // its stage input and accumulator start (and stay) zero there
    if (band) {
      const C<T> z = mkc<T>((T)0, (T)0);
      LD_TRY(fill<C<T>>(st, acc, 3 * CF, z));
      LD_TRY(fill<C<T>>(st, us, 3 * CF, z));
    }
    return LD_OK;
  }
  int state_get(void* ukout, void* nout, double* coord) override {
    LD_TRY(join(st, st2, evB));
    for (int b = 0; b < B; ++b) {
      if (ukout)
        LD_HIP(hipMemcpyAsync((C<T>*)ukout + (size_t)b * 3 * CF, at(uk, b), 3 * CF * sizeof(C<T>),
                              hipMemcpyDeviceToDevice, st));
      if (nout)
        LD_HIP(hipMemcpyAsync((T*)nout + (size_t)b * RF, at(n_state, b), RF * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    if (coord) LD_TRY(get_rows(coord));
    return LD_OK;
  }

  // Counting sort of the droplet slots by spatial bin of the base position
  // (on the density stream, which owns the droplet state during a step).
  int sort_droplets(hipStream_t s) {
    sorted_ok = false;
    if (np < 1 || !sort_enabled) return LD_OK;
    const unsigned pb = (unsigned)((np + 255) / 256);
    LD_TRY(fill<int>(s, hist, NB, 0));
    hipLaunchKernelGGL((k_bin_count<N>), dim3(pb, 1, B), dim3(256), 0, s, (const double*)x0, (int)np, (int)cap, p.L,
                       dp.invL, dp.inv_dx, pkey, prank, hist, rs);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_local, dim3(NBS, 1, B), dim3(1024), 0, s, hist, NB, bsum, rs);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_bsums, dim3(1, 1, B), dim3(1024), 0, s, bsum, NBS, rs);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_add, dim3((NB + 255) / 256, 1, B), dim3(256), 0, s, hist, NB, (const int*)bsum, rs);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_bin_scatter<PT>), dim3(pb, 1, B), dim3(256), 0, s, (int)np, (int)cap, (const int*)pkey,
                       (const int*)prank, (const int*)hist, (const double*)x0, (const PT*)v0, (const int64_t*)rid,
                       x0b, v0b, rid2, rs);
    LD_LAUNCHED();
    swap_sorted();
    return LD_OK;
  }
  // The sort scatters into the second set of droplet buffers; they become current.
  void swap_sorted() {
    std::swap(x0, x0b);
    std::swap(v0, v0b);
    std::swap(rid, rid2);
    sorted_ok = true;
    parity ^= 1;
  }
  bool sorts() const { return np >= 1 && sort_enabled; }

  // ------------------------------------------------------------ HIP graphs
  // With graphs on, one step (sort, four stages, epilogue on both streams) is
  // captured once per (h, droplet count, buffer parity) and replayed with one
  // hipGraphLaunch.  The capture runs the host side of the step (the sort's
  // buffer swap included) without executing it; a replay repeats that swap.
  // Steps run on an internal stream, joined to the caller's stream by events.
  struct StepGraph { hipGraphExec_t exec = nullptr; double h = 0.0; int64_t np = -1; bool band = false; };
  bool use_graph = false;
  hipStream_t stg = nullptr;
  hipStream_t stf = nullptr;          // flow stream with a CU mask (LD_CU_MASK), else unused
  hipEvent_t evG0 = nullptr, evG1 = nullptr;   // caller <-> internal stream joins
  StepGraph graphs[2];
  int parity = 0;                    // droplet buffers swapped an odd number of times

  int step_body(double h) {
    LD_TRY(sort_droplets(st2));
    LD_TRY(rk4_stages(h));
    return epilogue();
  }
  int graph_step(double h) {
    StepGraph& g = graphs[parity];
    if (!g.exec || g.h != h || g.np != np || g.band != band) {
      if (g.exec) { LD_HIP(hipGraphExecDestroy(g.exec)); g.exec = nullptr; }
      const int par = parity;
      LD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = join(st2, st, evA);                  // fork the density stream into the capture
      if (rc == LD_OK) rc = step_body(h);
      if (rc == LD_OK) rc = join(st, st2, evB);     // and join it back
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(st, &graph);
      if (rc != LD_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != LD_OK) return rc;
        LD_HIP(e);
      }
      const hipError_t ei = hipGraphInstantiate(&graphs[par].exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      LD_HIP(ei);
      graphs[par].h = h;
      graphs[par].np = np;
      graphs[par].band = band;
      LD_HIP(hipGraphLaunch(graphs[par].exec, st));
      return LD_OK;
    }
    LD_HIP(hipGraphLaunch(g.exec, st));
    if (sorts()) swap_sorted();
    return LD_OK;
  }

  int advance(double h, int nsteps) override {
    if (use_graph && prof_id < 0) {
      // steps on the internal stream (a legacy default stream cannot be captured)
      hipStream_t caller = st;
      LD_TRY(join(stg, caller, evG0));
      st = stg;
      int rc = LD_OK;
      for (int i = 0; i < nsteps && rc == LD_OK; ++i) rc = graph_step(h);
      st = caller;
      LD_TRY(rc);
      LD_TRY(join(caller, stg, evG1));
    } else if (stf) {
      hipStream_t caller = st;
      LD_TRY(join(stf, caller, evG0));
      st = stf;
      int rc = join(st2, st, evA);
      for (int i = 0; i < nsteps && rc == LD_OK; ++i) rc = step_body(h);
      if (rc == LD_OK) rc = join(st, st2, evB);
      st = caller;
      LD_TRY(rc);
      LD_TRY(join(caller, stf, evG1));
    } else {
      LD_TRY(join(st2, st, evA));      // state written on the caller's stream
      for (int i = 0; i < nsteps; ++i) LD_TRY(step_body(h));
      LD_TRY(join(st, st2, evB));      // caller's stream sees the whole step
    }
    std::vector<double> lm(B);
    std::vector<int32_t> bl(B);
    LD_TRY(status_all(lm.data(), bl.data()));
    for (int b = 0; b < B; ++b)
      if (bl[b]) { set_error("Threshold exceeded: max Re(u_hat) > 100 N^3 (realization %d)", b); return LD_ERR_BLOWUP; }
    return LD_OK;
  }
  // Largest max Re(u_hat) of the last epilogue and the blow-up flag, per
  // realization (host arrays of B entries).
  int status_all(double* last_max, int32_t* blown) override {
    LD_TRY(join(st, st2, evB));
    for (int b = 0; b < B; ++b) {
      if (last_max) LD_HIP(hipMemcpyAsync(last_max + b, at(lastmax, b), sizeof(double), hipMemcpyDeviceToHost, st));
      if (blown) LD_HIP(hipMemcpyAsync(blown + b, at(flag, b), sizeof(int), hipMemcpyDeviceToHost, st));
    }
    LD_HIP(hipStreamSynchronize(st));
    return LD_OK;
  }
  // Over the ensemble: the largest maximum and whether any realization blew up.
  int status(double* last_max, int32_t* blown) override {
    std::vector<double> lm(B);
    std::vector<int32_t> bl(B);
    LD_TRY(status_all(lm.data(), bl.data()));
    double m = lm[0];
    int32_t any = 0;
    for (int b = 0; b < B; ++b) { m = lm[b] > m ? lm[b] : m; any |= bl[b]; }
    if (last_max) *last_max = m;
    if (blown) *blown = any;
    return LD_OK;
  }

  int rk4(double h, const void* ukin, void* nio, const double* coord, int64_t n_, void* uknew, void* nnew,
          double* coord_new) override {
    LD_TRY(state_set(ukin, nio, coord, n_));
    LD_TRY(join(st2, st, evA));
    LD_TRY(rk4_stages(h));
    LD_TRY(join(st, st2, evB));
    LD_HIP(hipMemcpyAsync(uknew, acc, 3 * CF * sizeof(C<T>), hipMemcpyDeviceToDevice, st));
    LD_HIP(hipMemcpyAsync(nnew, nacc, RF * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (coord_new) LD_TRY(get_rows(coord_new));
    hipLaunchKernelGGL((k_relu<T>), dim3(1024), dim3(256), 0, st, (T*)nio, RF);   // in-place clip of n (:282 via :341)
    LD_LAUNCHED();
    return LD_OK;
  }

  // One RHS evaluation with the reference's full_RHS semantics (single stream).
  int full_rhs(const void* ukin, const void* nin, const double* coord, int64_t n_, void* ku, void* kn, double* kp,
               void* fcout, double visc, double forc, void* u_out, void* v_out) override {
    if (n_ > cap) { set_error("%lld droplets exceed the context capacity %lld", (long long)n_, (long long)cap); return LD_ERR_CAPACITY; }
    LD_TRY(join(st, st2, evB));
    const C<T>* ukp = (const C<T>*)ukin;
    T* uphys = u_out ? (T*)u_out : ubuf[0];
    T* vphys = v_out ? (T*)v_out : v;
    T* fcp = (T*)fcout;
    LD_TRY(shell(st, ukp, p.isforcing * forc));
    LD_TRY(nonlinear(st, ukp, uphys));
    RhsArgs<T> ra;
    ra.nl = w6; ra.us = ukp; ra.ubase = ukp; ra.acc = nullptr; ra.usnext = nullptr; ra.ku = (C<T>*)ku;
    ra.fac = fac; ra.fs = CF; ra.nu_eff = (T)(p.nu * p.isexplicit * visc); ra.tps = (T)p.tps;
    ra.wacc = (T)0; ra.anext = (T)0; ra.scale = inv(); ra.acc_init = 0;
    ra.kcut = kc; ra.band = 0;
    LD_TRY(xrhs(st, ra));
    // the physical-space v the droplets see uses visc as well (forced-dns-sm.py:319-321)
    LD_TRY(velocity(st, vphys, nullptr, (const T*)nin, clip_none, nullptr));
    LD_TRY(fill<T>(st, fcp, RF, (T)0));
    if (n_ > 0) {
      hipLaunchKernelGGL((k_prhs<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_,
                         (const T*)uphys, (const T*)vphys, (const T*)nin, (const double*)clip_none, fcp, kp, rows_a,
                         wrow, dp);
      LD_LAUNCHED();
    }
    OpZKn<T, N, P1> ok;
    ok.w = wn; ok.fs = CF; ok.scale = inv(); ok.kn_out = (T*)kn;
    ok.n0 = nullptr; ok.nacc = nullptr; ok.nnext = nullptr; ok.wacc = (T)0; ok.anext = (T)0; ok.acc_init = 0;
    ok.sum_target = -1; ok.partials = partials2;
    LD_TRY(density_rhs(st, fcp, &ok));
    return LD_OK;
  }

  // ------------------------------------------------------------ plain form
  // The reference's dataflow, unfused: every 3-D transform separate and one
  // kernel per array statement of forced-dns-sm.py:287-362 and 582-620, on
  // the caller's stream, droplets as (Np, 7) rows without reordering.
  T *pr3 = nullptr, *knp = nullptr;
  C<T> *kbuf = nullptr, *scr3 = nullptr;
  double *rows_c = nullptr, *rows_d = nullptr, *rows_e = nullptr, *rows_f = nullptr, *maxpart2 = nullptr;
  static constexpr int NBLK_SPEC = (int)((CF + 255) / 256);

  int plain_alloc() {
    if (pr3) return LD_OK;
    LD_TRY(alloc(pr3, 3 * RF)); LD_TRY(alloc(knp, RF)); LD_TRY(alloc(kbuf, 3 * CF)); LD_TRY(alloc(scr3, 3 * CF));
    LD_TRY(alloc(maxpart2, NBLK_SPEC));
    if (cap > 0) {
      LD_TRY(alloc(rows_c, 7 * cap)); LD_TRY(alloc(rows_d, 7 * cap));
      LD_TRY(alloc(rows_e, 7 * cap)); LD_TRY(alloc(rows_f, 7 * cap));
    }
    return LD_OK;
  }
  static dim3 g1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }
  // rfft_mpi: z, y, x passes (forced-dns-sm.py:201-206)
  int p_rfft(const T* in, C<T>* out, int nf) {
    for (int f = 0; f < nf; ++f) {
      OpZFwd<T, N, P1, 1> op{in + f * RF, out + f * CF, RF, CF};
      LD_TRY((zrun<P1, 0, 1>(st, op)));
    }
    LD_TRY((pass<1, -1>(st, out, out, nf, (T)1)));
    LD_TRY((pass<0, -1>(st, out, out, nf, (T)1)));
    return LD_OK;
  }
  // irfft_mpi: x, y, z passes (forced-dns-sm.py:208-213), nf <= 3
  int p_irfft(const C<T>* in, T* out, int nf) {
    LD_TRY((pass<0, +1>(st, in, scr3, nf, inv())));
    LD_TRY((pass<1, +1>(st, scr3, scr3, nf, inv())));
    for (int f = 0; f < nf; ++f) {
      OpZInv<T, N, P1, 1> op{scr3 + f * CF, out + f * RF, CF, RF, inv()};
      LD_TRY((zrun<P1, 1, 0>(st, op)));
    }
    return LD_OK;
  }
  template <class X, class S> int p_axpy(const X* x, const X* z, X* y, size_t n, S a) {
    hipLaunchKernelGGL((k_pl_axpy<X, S>), g1(n), dim3(256), 0, st, x, z, y, n, a);
    LD_LAUNCHED();
    return LD_OK;
  }
  int p_clip_zero(T* x, T* out) {   // clip_zero (:278-285): sums, clip x in place, out = x*old/new
    LD_TRY(clip_sums(st, x, clip_end));
    hipLaunchKernelGGL((k_clip_apply<T>), dim3(1024), dim3(256), 0, st, x, out, RF, (const double*)clip_end, 1, (size_t)0);
    LD_LAUNCHED();
    return LD_OK;
  }

  // full_RHS (:287-332) on stage inputs; ku -> kbuf, kn -> knp, kp -> rows_e.
  int plain_rhs(const C<T>* ukp, const T* nsp, const double* coord_s) {
    C<T>* fk = wn;
    T* u_ = ubuf[0];
    T* om = ubuf[1];
    LD_TRY(shell(st, ukp, p.isforcing));                                                         // :289
    hipLaunchKernelGGL((k_forcing_apply<T, N>), g1(CF), dim3(256), 0, st, ukp, fk, CF, (const double*)fac);
    LD_LAUNCHED();
    LD_TRY(p_irfft(ukp, u_, 3));                                                                  // :291-293
    hipLaunchKernelGGL((k_pl_curl<T, N>), g1(CF), dim3(256), 0, st, ukp, w6, CF);                  // :295-297
    LD_LAUNCHED();
    LD_TRY(p_irfft(w6, om, 3));
    hipLaunchKernelGGL((k_pl_cross<T>), g1(RF), dim3(256), 0, st, (const T*)u_, (const T*)om, pr3, RF);  // :299-301
    LD_LAUNCHED();
    C<T>* rk = w6 + 3 * CF;
    LD_TRY(p_rfft(pr3, rk, 3));                                                                    // :304-306
    hipLaunchKernelGGL((k_pl_add_mask<T, N>), g1(CF), dim3(256), 0, st, rk, (const C<T>*)fk, CF);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_pl_pressure<T, N>), g1(CF), dim3(256), 0, st, (const C<T>*)rk, ukp, kbuf, w6, CF,
                       (T)(p.nu * p.isexplicit), (T)p.tps);                                        // :309-321
    LD_LAUNCHED();
    LD_TRY(p_irfft(w6, v, 3));
    hipLaunchKernelGGL((k_pl_shift<T>), g1(RF), dim3(256), 0, st, v + 2 * RF, RF, (T)(p.tps * p.g));
    LD_LAUNCHED();
    LD_TRY(fill<T>(st, fc, RF, (T)0));
    if (np > 0) {                                                                                   // :324
      hipLaunchKernelGGL((k_prhs<T, N>), g1(np), dim3(256), 0, st, coord_s, (int)np, (const T*)u_, (const T*)v, nsp,
                         (const double*)clip_none, fc, rows_e, rows_a, wrow, dp);
      LD_LAUNCHED();
    }
    LD_TRY(p_rfft(fc, wn + 3 * CF, 1));                                                             // :325
    hipLaunchKernelGGL((k_pl_mask<T, N>), g1(CF), dim3(256), 0, st, wn + 3 * CF, CF, 1);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_pl_nv<T>), g1(RF), dim3(256), 0, st, nsp, (const T*)v, pr3, RF);          // :326
    LD_LAUNCHED();
    LD_TRY(p_rfft(pr3, w6, 3));                                                                      // :327-329
    hipLaunchKernelGGL((k_pl_mask<T, N>), g1(CF), dim3(256), 0, st, w6, CF, 3);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_pl_div<T, N>), g1(CF), dim3(256), 0, st, (const C<T>*)w6, (const C<T>*)(wn + 3 * CF), wn,
                       CF);                                                                          // :330
    LD_LAUNCHED();
    LD_TRY(p_irfft(wn, knp, 1));
    return LD_OK;
  }

  int advance_plain(double h, int nsteps) override {
    band = false;   // until the check at the end (an early error leaves it off)
    LD_TRY(plain_alloc());
    LD_TRY(join(st, st2, evB));
    const size_t nr = 7 * (size_t)np;
    if (np > 0) LD_TRY(get_rows(rows_c));        // droplet rows in input order
    static const double a_coef[4] = {0.0, 0.5, 0.5, 1.0};
    for (int it = 0; it < nsteps; ++it) {
      // RK4 (:335-362)
      if (np > 0) LD_HIP(hipMemcpyAsync(rows_d, rows_c, nr * sizeof(double), hipMemcpyDeviceToDevice, st));  // :337
      LD_HIP(hipMemcpyAsync(acc, uk, 3 * CF * sizeof(C<T>), hipMemcpyDeviceToDevice, st));                  // :338
      LD_HIP(hipMemcpyAsync(nacc, n_state, RF * sizeof(T), hipMemcpyDeviceToDevice, st));                   // :339
      for (int s = 0; s < 4; ++s) {
        const C<T>* ukp = uk;
        const double* cs = rows_c;
        if (s == 0) {
          LD_TRY(p_clip_zero(n_state, n_stage));                                                           // :341
        } else {
          const double a = a_coef[s] * h;
          LD_TRY(p_axpy(uk, (const C<T>*)kbuf, us, 3 * CF, (T)a));                                         // uk + ku*h/2
          LD_TRY(p_axpy(n_state, (const T*)knp, pr3, RF, (T)a));                                           // n + kn*h/2
          LD_TRY(p_clip_zero(pr3, n_stage));
          if (np > 0) LD_TRY(p_axpy(rows_c, (const double*)rows_e, rows_f, nr, a));                        // coord + h/2 kp
          ukp = us;
          cs = rows_f;
        }
        LD_TRY(plain_rhs(ukp, n_stage, cs));
        const double w = (s == 0 || s == 3) ? h / 6.0 : h / 3.0;
        if (np > 0) LD_TRY(p_axpy(rows_d, (const double*)rows_e, rows_d, nr, w));                          // sump += w kp
        LD_TRY(p_axpy(acc, (const C<T>*)kbuf, acc, 3 * CF, (T)w));                                         // uknew += w ku
        LD_TRY(p_axpy(nacc, (const T*)knp, nacc, RF, (T)w));                                               // nnew += w kn
      }
      // epilogue (:582-620)
      if (np > 0) LD_HIP(hipMemcpyAsync(rows_c, rows_d, nr * sizeof(double), hipMemcpyDeviceToDevice, st));  // :582
      LD_TRY(p_clip_zero(nacc, n_state));                                                                  // :584
      LD_TRY(p_irfft(acc, ubuf[0], 3));                                                                    // :597-599
      LD_TRY(p_rfft(ubuf[0], uk, 3));                                                                      // :601-603
      hipLaunchKernelGGL((k_pl_project<T, N>), dim3(NBLK_SPEC), dim3(256), 0, st, uk, CF, maxpart2);       // :607-610
      LD_LAUNCHED();
      hipLaunchKernelGGL(k_max_final_n, dim3(1), dim3(256), 0, st, (const double*)maxpart2, NBLK_SPEC,
                         100.0 * (double)N * N * N, flag, lastmax);                                         // :618
      LD_LAUNCHED();
    }
    if (np > 0) LD_TRY(set_rows(rows_c, np));
    LD_TRY(check_band());   // the transforms' round-off reaches every mode
    int32_t blown = 0;
    LD_HIP(hipMemcpyAsync(&blown, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    LD_HIP(hipStreamSynchronize(st));
    if (blown) { set_error("Threshold exceeded: max Re(u_hat) > 100 N^3"); return LD_ERR_BLOWUP; }
    return LD_OK;
  }

  // ------------------------------------------------------------ diagnostics
  double* dpart = nullptr;
  static constexpr int NBLK_DIAG = 256;
  int diag_alloc() {
    if (!dpart) LD_TRY(alloc(dpart, (size_t)NBLK_DIAG * (Nf > 4 ? Nf : 4)));
    return LD_OK;
  }
  double norm0() const { return 1.0 / ((double)N * N * N * N * N * N / std::pow(2.0 * M_PI, 3.0)); }
  int spectrum(const void* f, int nf, double w, double* out) override {
    LD_TRY(diag_alloc());
    return spectrum_ws(f, nf, w, out, dpart, (size_t)NBLK_DIAG * (Nf > 4 ? Nf : 4));
  }
  // The shell partial sums go to ws (every entry the final sum reads is
  // written first, so ws needs no clearing); nothing is allocated.
  int spectrum_ws(const void* f, int nf, double w, double* out, double* ws, size_t ws_len) override {
    if ((size_t)NBLK_DIAG * Nf > ws_len) {
      set_error("spectrum scratch of %zu doubles is below the %zu this grid needs", ws_len, (size_t)NBLK_DIAG * Nf);
      return LD_ERR_ARG;
    }
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_shell_partial<T, N>), dim3(NBLK_DIAG), dim3(256), 0, st, (const C<T>*)f, (const C<T>*)nullptr,
                       nf, CF, 0, w, norm0(), ws);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_shell_final<N>), dim3(1), dim3(256), 0, st, (const double*)ws, NBLK_DIAG, 0, out);
    LD_LAUNCHED();
    return LD_OK;
  }
  int flux(const void* ukin, const void* kuin, double* out) override {
    LD_TRY(diag_alloc());
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_shell_partial<T, N>), dim3(NBLK_DIAG), dim3(256), 0, st, (const C<T>*)ukin,
                       (const C<T>*)kuin, 3, CF, 1, 1.0, norm0(), dpart);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_shell_final<N>), dim3(1), dim3(256), 0, st, (const double*)dpart, NBLK_DIAG, 1, out);
    LD_LAUNCHED();
    return LD_OK;
  }
  // Dissipation spectrum nu |k|^2 sum_c |u_c|^2 normalize by shell; its sum is
  // the dissipation rate 2 nu sum_k k^2 E(k).
  int dissipation(const void* ukin, double* out) override {
    LD_TRY(diag_alloc());
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_shell_partial<T, N>), dim3(NBLK_DIAG), dim3(256), 0, st, (const C<T>*)ukin,
                       (const C<T>*)nullptr, 3, CF, 2, p.nu, norm0(), dpart);
    LD_LAUNCHED();
    hipLaunchKernelGGL((k_shell_final<N>), dim3(1), dim3(256), 0, st, (const double*)dpart, NBLK_DIAG, 0, out);
    LD_LAUNCHED();
    return LD_OK;
  }
  // Physical divergence field div = d_x u_x + d_y u_y + d_z u_z of the
  // spectral velocity uk (forced-dns-sm.py:216-229, 472); wn is its spectral
  // work buffer, w6 the transform's.
  int divergence(const void* ukin, void* divout) override {
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_div_spec<T, N>), dim3(1024), dim3(256), 0, st, (const C<T>*)ukin, CF, wn);
    LD_LAUNCHED();
    return irfft(wn, divout, 1);
  }
  // Sum, min, max and sum of squares of the resident droplets' mass column
  // (m/M0; plot.py:56-72 tracks the total).
  int droplet_stats(double* out) override {
    LD_TRY(diag_alloc());
    LD_TRY(join(st, st2, evB));
    if (np == 0) {
      const double z[4] = {0.0, 0.0, 0.0, 0.0};
      LD_HIP(hipMemcpyAsync(out, z, sizeof(z), hipMemcpyHostToDevice, st));
      LD_HIP(hipStreamSynchronize(st));
      return LD_OK;
    }
    hipLaunchKernelGGL((k_stats_partial<PT>), dim3(NBLK_DIAG), dim3(256), 0, st, (const PT*)(v0 + 3 * cap),
                       (size_t)np, dpart);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, st, (const double*)dpart, NBLK_DIAG, out);
    LD_LAUNCHED();
    return LD_OK;
  }
  int stats(const void* x, int nf, double* out) override {
    LD_TRY(diag_alloc());
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_stats_partial<T>), dim3(NBLK_DIAG), dim3(256), 0, st, (const T*)x, (size_t)nf * RF, dpart);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, st, (const double*)dpart, NBLK_DIAG, out);
    LD_LAUNCHED();
    return LD_OK;
  }

  // ------------------------------------------------------------ collisions
  int *ckey = nullptr, *crank = nullptr, *corder = nullptr, *chist = nullptr, *cbsum = nullptr, *cused = nullptr;
  unsigned long long *ckeys = nullptr, *ccount = nullptr;
  int *cnev = nullptr;
  int64_t ccap = 0, ckcap = 0;
  int chist_cap = 0;

  static int64_t pow2ceil(int64_t x) { int64_t v = 2; while (v < x) v <<= 1; return v; }

  // Cell list of the droplet rows on st: cells of side >= reach (fewer than 3
  // per side -> one cell, all pairs); bin starts in chist, slots in corder.
  int build_cells(const double* rows, int64_t n_, double reach, int& nc, double& cell) {
    nc = (reach > 0.0) ? (int)floor(p.L / reach) : 256;
    if (nc > 256) nc = 256;
    if (nc < 3) nc = 1;
    cell = p.L / nc;
    const int ncells = nc * nc * nc;
    if (n_ > ccap) {
      LD_TRY(alloc(ckey, n_)); LD_TRY(alloc(crank, n_)); LD_TRY(alloc(corder, n_)); LD_TRY(alloc(cused, n_));
      ccap = n_;
    }
    if (ncells > chist_cap) {
      LD_TRY(alloc(chist, ncells)); LD_TRY(alloc(cbsum, (ncells + 4095) / 4096));
      chist_cap = ncells;
    }
    const unsigned pb = (unsigned)((n_ + 255) / 256);
    LD_TRY(fill<int>(st, chist, ncells, 0));
    hipLaunchKernelGGL(k_coll_count, dim3(pb), dim3(256), 0, st, rows, (int)n_, p.L, cell, nc, ckey, crank, chist);
    LD_LAUNCHED();
    const int nbs = (ncells + 4095) / 4096;
    hipLaunchKernelGGL(k_scan_local, dim3(nbs), dim3(1024), 0, st, chist, ncells, cbsum, (size_t)0);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(1024), 0, st, cbsum, nbs, (size_t)0);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_scan_add, dim3((ncells + 255) / 256), dim3(256), 0, st, chist, ncells, (const int*)cbsum,
                       (size_t)0);
    LD_LAUNCHED();
    hipLaunchKernelGGL(k_coll_scatter, dim3(pb), dim3(256), 0, st, (int)n_, (const int*)ckey, (const int*)crank,
                       (const int*)chist, corder);
    LD_LAUNCHED();
    return LD_OK;
  }

  int pair_stats(const double* rows, int64_t n_, double rmax, int nbins, int64_t* counts, double* wsum,
                 double* wneg) override {
    if (nbins < 1 || nbins > 512) { set_error("nbins must be in [1, 512]"); return LD_ERR_ARG; }
    if (!(rmax > 0.0) || rmax > 0.5 * p.L) { set_error("rmax must be in (0, L/2]"); return LD_ERR_ARG; }
    if (n_ > 0x7fffffff) { set_error("too many droplets"); return LD_ERR_ARG; }
    LD_TRY(join(st, st2, evB));
    LD_TRY(fill<unsigned long long>(st, (unsigned long long*)counts, (size_t)nbins, 0ull));
    LD_TRY(fill<double>(st, wsum, (size_t)nbins, 0.0));
    LD_TRY(fill<double>(st, wneg, (size_t)nbins, 0.0));
    if (n_ < 2) return LD_OK;
    int nc;
    double cell;
    LD_TRY(build_cells(rows, n_, rmax, nc, cell));
    hipLaunchKernelGGL(k_pair_stats, dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, rows, (int)n_, p.L, cell,
                       nc, (const int*)chist, (const int*)corder, (const int*)ckey, rmax, rmax / nbins, nbins,
                       (unsigned long long*)counts, wsum, wneg);
    LD_LAUNCHED();
    return LD_OK;
  }

  int collide(const double* rows, const double* rad, int64_t n_, double rmax, int64_t* pairs, int64_t max_pairs,
              int64_t* count) override {
    *count = 0;
    if (n_ < 2) return LD_OK;
    if (n_ > 0x7fffffff) { set_error("too many droplets"); return LD_ERR_ARG; }
    LD_TRY(join(st, st2, evB));
    // cells of side >= 2 rmax (the largest pair distance)
    int nc;
    double cell;
    LD_TRY(build_cells(rows, n_, 2.0 * rmax, nc, cell));
    const int64_t kc = pow2ceil(max_pairs > 2 ? max_pairs : 2);
    if (kc > ckcap) { LD_TRY(alloc(ckeys, kc)); ckcap = kc; }
    if (!ccount) { LD_TRY(alloc(ccount, 1)); LD_TRY(alloc(cnev, 1)); }
    const unsigned pb = (unsigned)((n_ + 255) / 256);
    LD_TRY(fill<unsigned long long>(st, ccount, 1, 0ull));
    hipLaunchKernelGGL(k_coll_pairs, dim3(pb), dim3(256), 0, st, rows, rad, (int)n_, p.L, cell, nc,
                       (const int*)chist, (const int*)corder, (const int*)ckey, ckeys,
                       (unsigned long long)max_pairs, ccount);
    LD_LAUNCHED();
    unsigned long long cnt = 0;
    LD_HIP(hipMemcpyAsync(&cnt, ccount, sizeof(cnt), hipMemcpyDeviceToHost, st));
    LD_HIP(hipStreamSynchronize(st));
    *count = (int64_t)cnt;
    if ((int64_t)cnt > max_pairs) {
      set_error("%llu collision pairs exceed max_pairs = %lld", cnt, (long long)max_pairs);
      return LD_ERR_CAPACITY;
    }
    if (cnt == 0) return LD_OK;
    const int n2 = (int)pow2ceil((int64_t)cnt);
    if ((int64_t)n2 > (int64_t)cnt)
      LD_TRY(fill<unsigned long long>(st, ckeys + cnt, (size_t)(n2 - (int64_t)cnt), ~0ull));
    for (int k = 2; k <= n2; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        hipLaunchKernelGGL(k_bitonic, dim3((n2 + 255) / 256), dim3(256), 0, st, ckeys, n2, j, k);
        LD_LAUNCHED();
      }
    hipLaunchKernelGGL(k_pairs_out, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                       (const unsigned long long*)ckeys, (int)cnt, pairs);
    LD_LAUNCHED();
    return LD_OK;
  }

  int coalesce(double* rows, int64_t n_, const int64_t* pairs, int64_t npairs, int64_t* events,
               int64_t* nevents) override {
    *nevents = 0;
    if (npairs <= 0) return LD_OK;
    if (n_ > ccap) {
      LD_TRY(alloc(ckey, n_)); LD_TRY(alloc(crank, n_)); LD_TRY(alloc(corder, n_)); LD_TRY(alloc(cused, n_));
      ccap = n_;
    }
    if (!ccount) { LD_TRY(alloc(ccount, 1)); LD_TRY(alloc(cnev, 1)); }
    LD_TRY(join(st, st2, evB));
    LD_TRY(fill<int>(st, cused, (size_t)n_, 0));
    hipLaunchKernelGGL(k_coalesce, dim3(1), dim3(64), 0, st, rows, pairs, (int)npairs, p.L, cused, events, cnev);
    LD_LAUNCHED();
    int ne = 0;
    LD_HIP(hipMemcpyAsync(&ne, cnev, sizeof(int), hipMemcpyDeviceToHost, st));
    LD_HIP(hipStreamSynchronize(st));
    *nevents = ne;
    return LD_OK;
  }

  // ------------------------------------------------------------ droplet API
  int interp(const double* coord, int64_t n_, const void* f, int nc, double* out) override {
    if (n_ <= 0) return LD_OK;
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_interp<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_,
                       (const T*)f, nc, out, dp);
    LD_LAUNCHED();
    return LD_OK;
  }
  int exterp(const double* coord, int64_t n_, const double* w, void* c) override {
    LD_TRY(join(st, st2, evB));
    LD_TRY(fill<T>(st, (T*)c, RF, (T)0));
    if (n_ <= 0) return LD_OK;
    hipLaunchKernelGGL((k_exterp<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_, w,
                       (T*)c, dp);
    LD_LAUNCHED();
    return LD_OK;
  }
  // ------------------------------------------------------------ learned subgrid form
  int resample(const void* src, int ns, void* dst, int nf) override {
    LD_TRY(join(st, st2, evB));
    const double sc = std::pow((double)N / (double)ns, 3.0);
    hipLaunchKernelGGL((k_resample<T, N>), dim3(2048), dim3(256), 0, st, (const C<T>*)src, ns, (C<T>*)dst, nf, (T)sc);
    LD_LAUNCHED();
    return LD_OK;
  }
  int grad_spec(const void* ukin, void* gk, double s) override {
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_grad_spec<T, N>), dim3(2048), dim3(256), 0, st, (const C<T>*)ukin, (C<T>*)gk, (T)s);
    LD_LAUNCHED();
    return LD_OK;
  }
  int fields_to_rows(const void* f, int nc, float* rows, float s) override {
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_fields_to_rows<T, N>), dim3(2048), dim3(256), 0, st, (const T*)f, nc, rows, s);
    LD_LAUNCHED();
    return LD_OK;
  }
  int subgrid(const float* net_in, float* net_out, int r, const void* base, const void* fine, void* out,
              float s) override {
    LD_TRY(join(st, st2, evB));
    hipLaunchKernelGGL((k_subgrid<T, N>), dim3(2048), dim3(256), 0, st, net_in, net_out, r, (const T*)base,
                       (const T*)fine, (T*)out, s);
    LD_LAUNCHED();
    return LD_OK;
  }
  int frozen_rk4(double* rows, int64_t n_, const void* u, double h, int nsteps) override {
    if (n_ <= 0 || nsteps <= 0) return LD_OK;
    LD_TRY(join(st, st2, evB));
    // the step's gather records are free between steps
    V4<T>* g = grec[0];
    hipLaunchKernelGGL((k_u_records<T, N>), dim3(2048), dim3(256), 0, st, (const T*)u, g);
    LD_LAUNCHED();
    // 4 lanes per droplet; block count a multiple of 8 for the XCD mapping
    const unsigned nb = (unsigned)(((4 * n_ + 255) / 256 + 7) / 8 * 8);
    hipLaunchKernelGGL((k_frozen_rk4<T, N>), dim3(nb), dim3(256), 0, st, rows, (int)n_, (const V4<T>*)g, h, nsteps,
                       dp);
    LD_LAUNCHED();
    return LD_OK;
  }
  int prhs(const double* coord, int64_t n_, const void* uf, const void* usf, const void* nf, void* fcout,
           double* rhs, double* interpm, double* exterpm) override {
    LD_TRY(join(st, st2, evB));
    LD_TRY(fill<T>(st, (T*)fcout, RF, (T)0));
    if (n_ <= 0) return LD_OK;
    hipLaunchKernelGGL((k_prhs<T, N>), dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, st, coord, (int)n_,
                       (const T*)uf, (const T*)usf, (const T*)nf, (const double*)clip_none, (T*)fcout, rhs, interpm,
                       exterpm, dp);
    LD_LAUNCHED();
    return LD_OK;
  }
};

}  // namespace ld
