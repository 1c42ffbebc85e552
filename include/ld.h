/*
 * ld.h — C-ABI of the MI355X forced-DNS + droplet step library (libld.so).
 *
 * Drop-in boundary for the hot path of the reference (Rajarshi-prime/
 * lucky-droplets): forced-dns-sm.py (spectral flow, RK4, step epilogue) and
 * particles.py (droplets).  The reference has no FFI of its own: its hot path
 * is a set of Python functions on module-level numpy arrays.  Each entry
 * point below replaces one of those functions; the comment names the
 * reference function (file:line) and the argument meaning is kept:
 * caller-owned buffers are written in place, as the reference mutates its
 * preallocated arrays.
 *
 * Conventions
 *  - All buffer arguments are DEVICE pointers (e.g. torch tensor.data_ptr()).
 *    Field arrays (physical and spectral) must be 32-byte aligned (they are
 *    moved as 16/32-byte vectors; torch allocations are 256-byte aligned);
 *    a misaligned field pointer returns LD_ERR_ARG.
 *  - Field precision follows the context: LD_F32 -> float / complex64,
 *    LD_F64 -> double / complex128.  Droplet rows are always float64.
 *  - Layouts are the reference's at one MPI rank (C order):
 *      physical field  (N, N, N)          [x][y][z]
 *      spectral field  (N, N, N/2+1)      [kx][ky][kz], complex
 *      vector field    3 fields back to back, component-major
 *      droplet rows    (Np, 7) float64    [x, y, z, vx, vy, vz, m/M0]
 *  - Forward transforms are unnormalised; inverse transforms carry 1/N^3.
 *  - Every call is asynchronous on the context's stream (ld_set_stream),
 *    except where noted: ld_state_set (and ld_rk4 / ld_step, which set the
 *    state) waits once for its check of the spectral state, and the stepping
 *    calls wait to read the blow-up flag.
 *  - Return value: LD_OK or an LD_ERR_* code; ld_last_error() gives text.
 *  - Contexts are independent; a context must not be used from two threads
 *    at once.
 */
#ifndef LD_H
#define LD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LD_OK 0
#define LD_ERR_ARG 1        /* bad argument (size, null pointer, unsupported N) */
#define LD_ERR_HIP 2        /* HIP runtime error */
#define LD_ERR_BLOWUP 3     /* max Re(u_hat) > 100 N^3 (forced-dns-sm.py:618-620) */
#define LD_ERR_CAPACITY 4   /* more droplets than the context was created for */

#define LD_F32 0
#define LD_F64 1

/* Physical and numerical parameters: the reference's module constants. */
typedef struct ld_params {
  int32_t N;             /* grid points per side; power of two, 16..512       forced-dns-sm.py:28 */
  int32_t precision;     /* LD_F32 or LD_F64 (fields and transforms) */
  double L;              /* box side, 2*pi                                    forced-dns-sm.py:54 */
  double nu;             /* viscosity                                         forced-dns-sm.py:83 */
  int32_t lp;            /* viscosity power; only 1                           forced-dns-sm.py:78 */
  double isexplicit;     /* 1.0: explicit viscosity                           forced-dns-sm.py:23 */
  int32_t nshells;       /* number of forced shells (<= 8)                    forced-dns-sm.py:89 */
  int32_t shell_no[8];   /* forced shells, each in [1, 8]                     forced-dns-sm.py:90 */
  double f0;             /* power input per shell                             forced-dns-sm.py:97 */
  double isforcing;      /* 1.0 forcing on                                    forced-dns-sm.py:19 */
  double tps;            /* small-droplet relaxation time tau_p = sts*tf      forced-dns-sm.py:93 */
  double g;              /* gravity in simulation units                       forced-dns-sm.py:94 */
  double st_s;           /* small-droplet Stokes number (sts)                 particles.py:95 */
  double tau_eta;        /* Kolmogorov time used by the droplets (tf)         particles.py:56 */
  double factor;         /* m/M0 = St^1.5 * factor                            particles.py:74 */
  double growthfactor;   /* 9 pi nu / (2 rho_p)                               particles.py:75 */
  int64_t max_particles; /* droplet capacity of the context (per realization) */
  int32_t batch;         /* realizations of an ensemble (0 or 1: one).  With batch > 1
                            only the resident-state entry points (ld_state_*, ld_advance,
                            ld_status*) accept the context; no counterpart in the
                            reference, which runs one realization per job */
  int32_t use_graph;     /* 1: ld_advance captures one step into a HIP graph (per step
                            size and droplet count) and replays it; same results */
  int32_t autotune;      /* 1: ld_create times the launch-geometry variants of the y
                            passes and keeps the fastest (ld_tuning); same results */
} ld_params;

typedef struct ld_ctx ld_ctx;

/* Context: plans, twiddle tables and all workspace are allocated here, never
 * inside a step.  Replaces the module-level setup of forced-dns-sm.py:49-194
 * and MPI_particles.__init__/to_interp/to_exterp (particles.py:54-142). */
int ld_create(const ld_params* params, ld_ctx** out);
int ld_destroy(ld_ctx* ctx);
int ld_set_stream(ld_ctx* ctx, void* hip_stream);
/* Copies the calling thread's last error message (NUL-terminated). */
int ld_last_error(char* buf, size_t len);
/* Library version string, e.g. "ld 0.1 gfx950". */
const char* ld_version(void);

/* ---------------------------------------------------------------- flow */
/* rfft_mpi(u, fu): 3-D r2c of nfields real fields        forced-dns-sm.py:201-206 */
int ld_rfft(ld_ctx* ctx, const void* u, void* fu, int nfields);
/* irfft_mpi(fu, u): 3-D c2r of nfields fields, 1/N^3     forced-dns-sm.py:208-213 */
int ld_irfft(ld_ctx* ctx, const void* fu, void* u, int nfields);
/* forcing(uk, fk): constant-power solenoidal forcing     forced-dns-sm.py:239-275
 * ek_out (may be NULL): device double[N/2+1] shell energies of the forced shells. */
int ld_forcing(ld_ctx* ctx, const void* uk, void* fk, double* ek_out);
/* clip_zero(x): x clipped in place, out = x*oldmean/newmean  forced-dns-sm.py:278-285 */
int ld_clip_zero(ld_ctx* ctx, void* x, void* out);
/* full_RHS(t, uk, n, sump, coord, stb, ku, kn, fc, visc, forc)   forced-dns-sm.py:287-332
 * coord: droplet stage rows (np, 7); kp: droplet RHS rows (np, 7); fc: deposit field.
 * sump is passed through unchanged (rows keep their order on one device).
 * u_out, v_out (may be NULL): the physical u and slow-manifold v (3 fields each). */
int ld_full_rhs(ld_ctx* ctx, const void* uk, const void* n, const double* coord, int64_t np,
                void* ku, void* kn, double* kp, void* fc, double visc, double forc,
                void* u_out, void* v_out);
/* RK4(t, h, stb, uk, n, uknew, nnew) -> (sump, uknew, nnew)  forced-dns-sm.py:335-362
 * n is clipped in place as the reference's first clip_zero does (:341).
 * coord: droplet rows in (np, 7); coord_new: rows out (sump). */
int ld_rk4(ld_ctx* ctx, double h, const void* uk, void* n, const double* coord, int64_t np,
           void* uknew, void* nnew, double* coord_new);
/* Loop body of evolve_and_save, nsteps times            forced-dns-sm.py:582-620
 * (RK4, update_intrinsic, clip_zero, reality round trip, projection, blow-up
 * check) on uk (3 spectral fields), n (physical) and droplet rows, in place.
 * Returns LD_ERR_BLOWUP if any step's max Re(uk) exceeded 100 N^3. */
int ld_step(ld_ctx* ctx, double h, void* uk, void* n, double* coord, int64_t np, int32_t nsteps);

/* Context-resident state (the fast path: no layout conversion per step).
 * With batch B > 1 every array holds B realizations back to back: uk
 * (B, 3, N, N, N/2+1), n (B, N, N, N), coord (B, np, 7) with np rows per
 * realization; ld_advance steps all of them with one launch per kernel
 * (blockIdx.z = realization).  Each realization evolves exactly as it would
 * alone (same kernels, same arithmetic).  ld_state_set also checks whether
 * uk has any nonzero mode outside the dealiasing mask (one synchronisation);
 * without one, the steps skip the work whose inputs are zero there. */
int ld_state_set(ld_ctx* ctx, const void* uk, const void* n, const double* coord, int64_t np);
int ld_state_get(ld_ctx* ctx, void* uk, void* n, double* coord);
int ld_advance(ld_ctx* ctx, double h, int32_t nsteps);
/* The same steps with the reference's dataflow, unfused ("plain" form): every
 * 3-D transform separate (74 per step) and one kernel per array statement of
 * RK4/full_RHS and the loop body (forced-dns-sm.py:287-362, 582-620); droplet
 * rows are processed in input order.  The baseline of the fused path. */
int ld_advance_plain(ld_ctx* ctx, double h, int32_t nsteps);
/* Last step's max Re(uk) (host double) and sticky blow-up flag; over an
 * ensemble the largest maximum and whether any realization blew up. */
int ld_status(ld_ctx* ctx, double* last_max, int32_t* blown_up);
/* The same per realization: last_max and blown_up are host arrays of batch
 * entries.  No counterpart in the reference (one realization per job). */
int ld_status_all(ld_ctx* ctx, double* last_max, int32_t* blown_up);
/* The autotuner's choice: lines per block of the y passes (*ylines, host) and
 * the measured median times in microseconds of the candidates (us[3], host:
 * M, M/2 and 2M lines with M = 16 for FP32 and 8 for FP64; 0 for a candidate
 * not timed, all zeros when the context was created without autotune).  No
 * counterpart in the reference. */
int ld_tuning(ld_ctx* ctx, int32_t* ylines, double* us);

/* Every launch-geometry choice (host arrays): choice[7] = y-pass lines per
 * block, z-tile height of the u x omega pass and of the v pass, lines per
 * block of the u/omega, RHS and density x passes, and the threads per line N1
 * of the y passes' line split N = N1 N2; us[19] = the y-pass candidates'
 * median times (as ld_tuning), then per tuned kernel the times of its
 * default, its half (height or lines) and its double (z tiles only) variant
 * (0 where not timed), then the 6-field y-inverse at the chosen lines with
 * the other split, N1/2 threads per line.  Every choice the autotuner makes
 * gives the same results; the split is not one of them (another
 * factorisation rounds differently): it is timed for the record and set by
 * the environment (LD_YSPLIT=1).  No counterpart in the reference. */
int ld_tuning_all(ld_ctx* ctx, int32_t* choice, double* us);

/* ---------------------------------------------------------- measurement */
/* Kernel classes that ld_profile can time (HIP events recorded on the stream
 * each launch runs on).  No counterpart in the reference, which only prints
 * wall time per step (forced-dns-sm.py:565-577). */
#define LD_K_DROPLETS 0   /* fused droplet RK stage: gather, RHS, deposit input */
#define LD_K_XRHS 1       /* x pass: nonlinear term, forcing, projection, RK update, v_hat */
#define LD_K_XUW 2        /* x pass: u_hat, i k x u_hat */
#define LD_K_YPASS 3      /* y line passes (all) */
#define LD_K_ZCROSS 4     /* z pass: u, omega, u x omega */
#define LD_K_ZVEL 5       /* z pass: v, n v */
#define LD_K_XKN 6        /* x pass: density flux */
#define LD_K_ZKN 7        /* z pass: kn, RK update of n */
#define LD_K_DEPOSIT 8    /* droplet deposit: brick pass and fallback accumulate/fold */
/* Start timing every launch of kernel class `kernel` (or stop, enable = 0). */
int ld_profile(ld_ctx* ctx, int32_t kernel, int32_t enable);
/* Synchronise, then report the summed duration (ms) and launch count since
 * ld_profile was enabled; resets the counters. */
int ld_profile_read(ld_ctx* ctx, double* total_ms, int64_t* launches);

/* ------------------------------------------------------------ droplets */
/* MPI_particles.interp_cosine(coord, u): 4x4x4 cosine-kernel gather of ncomp
 * physical fields at the rows' positions -> out (np, ncomp)   particles.py:435-511 */
int ld_interp_cosine(ld_ctx* ctx, const double* coord, int64_t np, const void* fields, int32_t ncomp,
                     double* out);
/* MPI_particles.exterp_cosine_scalar(coord, c): c = deposit of w / dx^3
 * (c is zeroed first, particles.py:521)                        particles.py:514-577 */
int ld_exterp_cosine_scalar(ld_ctx* ctx, const double* coord, int64_t np, const double* w, void* c);
/* MPI_particles.pRHS(t, coord, u, us, c, fc, sump) -> (sump, rhs, fc)  particles.py:817-837
 * Also returns interpmat (np, 7) and exterpmat (np).  n is used as given. */
int ld_prhs(ld_ctx* ctx, const double* coord, int64_t np, const void* u, const void* us, const void* n,
            void* fc, double* rhs, double* interpmat, double* exterpmat);

/* --------------------------------------------------------- diagnostics */
/* Shell spectrum e3d_to_e1d(weight * sum_c |f_c|^2 * normalize) of nfields
 * spectral fields into out[N/2+1] (device double): weight 0.5 on u_hat gives
 * E(k) (forced-dns-sm.py:423, 431); weight 1 on n_hat the n spectrum (plot.py:37-38). */
int ld_spectrum(ld_ctx* ctx, const void* f, int32_t nfields, double weight, double* out);
/* ld_spectrum without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  No counterpart
 * in the reference.  Same arguments and results (bit for bit) as ld_spectrum;
 * its per-block partial sums live in a fixed 526,336-byte device array of the
 * library (256 blocks x N/2+1 shells, N <= 512), one per GPU, looked up at
 * ld_create.  Calls on one GPU share that array: calls that overlap in time
 * on one GPU race and must be serialised by the caller. */
int ld_spectrum_noalloc(ld_ctx* ctx, const void* f, int32_t nfields, double weight, double* out);
/* Energy flux Pi(k): e3d_to_e1d(Re(conj(uk) . ku) * dealias * normalize),
 * cumulated from the highest shell down (forced-dns-sm.py:426-428); ku from
 * ld_full_rhs with visc = 0, forc = 0 (:425).  out[N/2+1] device double. */
int ld_flux(ld_ctx* ctx, const void* uk, const void* ku, double* out);
/* Dissipation spectrum: e3d_to_e1d(nu |k|^2 sum_c |u_c|^2 * normalize) of the
 * spectral velocity uk into out[N/2+1] (device double); its sum is the
 * dissipation rate 2 nu sum_k k^2 E(k).  The reference leaves its dissipation
 * diagnostic commented out (forced-dns-sm.py:473); this is the standard one. */
int ld_dissipation(ld_ctx* ctx, const void* uk, double* out);
/* Physical divergence field div = diff_x(u) + diff_y(v) + diff_z(w) of the
 * spectral velocity uk (one real field out): the reference's spectral
 * derivatives forced-dns-sm.py:216-229, whose max |div| save() prints (:472). */
int ld_divergence(ld_ctx* ctx, const void* uk, void* div);
/* Droplets of the resident state (ld_state_set / ld_advance): sum, min, max
 * and sum of squares of the mass column m/M0 -> out[4] (device double); the
 * total droplet mass that plot.py:56-72 tracks over a run. */
int ld_droplet_stats(ld_ctx* ctx, double* out);
/* sum, min, max and sum of squares over nfields real fields -> out[4]
 * (device double): energy, n mean/min/max (forced-dns-sm.py:467-471). */
int ld_field_stats(ld_ctx* ctx, const void* x, int32_t nfields, double* out);

/* ---------------------------------------------------------- collisions */
/* Droplet-droplet collision pairs with a cell list.  No counterpart in the
 * reference (its droplets grow only by collecting the continuum density,
 * particles.py:828).  Pairs (i, j), i < j, of rows whose minimum-image FP64
 * distance satisfies d^2 < (r_i + r_j)^2 (d2 = (dx*dx + dy*dy) + dz*dz, no
 * fused multiply-add), sorted by (i, j), written to pairs (max_pairs, 2) int64.
 * rmax >= every radius.  *npairs (HOST) receives the count; LD_ERR_CAPACITY
 * if it exceeds max_pairs. */
int ld_collide_pairs(ld_ctx* ctx, const double* rows, const double* radius, int64_t np, double rmax,
                     int64_t* pairs, int64_t max_pairs, int64_t* npairs);
/* Pair statistics within the cutoff rmax (0 < rmax <= L/2), the input of the
 * radial distribution function and the kinematic collision kernel; no
 * counterpart in the reference.  For every pair i < j with minimum-image
 * d2 < rmax^2: dist = sqrt(d2), bin k = min((int)(dist / (rmax / nbins)),
 * nbins - 1), w_r = ((dvx*dx + dvy*dy) + dvz*dz) / dist (dv = v_j - v_i,
 * 0 if dist == 0); counts[k] += 1, wsum[k] += |w_r|, wneg[k] += max(-w_r, 0).
 * counts (nbins int64), wsum, wneg (nbins double): DEVICE, overwritten.
 * 1 <= nbins <= 512. */
int ld_pair_stats(ld_ctx* ctx, const double* rows, int64_t np, double rmax, int32_t nbins, int64_t* counts,
                  double* wsum, double* wneg);
/* Coalescence events for sorted pairs: visited in order, a pair merges only if
 * neither droplet merged earlier in this call; survivor i takes the summed
 * mass, mass-weighted velocity and position x_i + (m_j/m) d; row j is left
 * with zero mass and velocity.  events (npairs, 2) int64; *nevents (HOST). */
int ld_coalesce(ld_ctx* ctx, double* rows, int64_t np, const int64_t* pairs, int64_t npairs, int64_t* events,
                int64_t* nevents);

/* ---------------------------------------------------------- learned subgrid form */
/* A periodic 3-D CNN on an N^3 grid: nlayers 3x3x3 convolutions with bias,
 * ReLU between them, channels[0] inputs, channels[nlayers] outputs (each
 * 1..32), hidden widths 16 or 32; bf16 weights and activations, FP32
 * accumulation (MFMA), FP32 master parameters trained with Adam on a
 * mean-squared-error loss.  Fields are rows x[node][c], node = (i N + j) N + k,
 * FP32, device memory.  Parameter layout (FP32, device): per layer
 * W[tap][co][ci] (tap = (dx+1) 9 + (dy+1) 3 + (dz+1), offsets -1..1, channel
 * counts padded to 16 or 32, padded entries kept 0) then b[co].  No
 * counterpart in the reference. */
#define LD_CNN_MAX_LAYERS 6
typedef struct ld_cnn_params {
  int32_t N;                                  /* grid points per side, power of two, 4..512 */
  int32_t nlayers;                            /* 1..LD_CNN_MAX_LAYERS */
  int32_t channels[LD_CNN_MAX_LAYERS + 1];    /* channel count at each layer boundary */
} ld_cnn_params;
typedef struct ld_cnn ld_cnn;
int ld_cnn_create(const ld_cnn_params* p, ld_cnn** out);
int ld_cnn_destroy(ld_cnn* c);
int ld_cnn_set_stream(ld_cnn* c, void* hip_stream);
int ld_cnn_nparams(ld_cnn* c, int64_t* n);                 /* n: host */
int ld_cnn_set_params(ld_cnn* c, const float* params);     /* also resets the Adam moments */
int ld_cnn_get_params(ld_cnn* c, float* params);
/* y[node][channels[nlayers]] = net(x[node][channels[0]]) */
int ld_cnn_forward(ld_cnn* c, const float* x, float* y);
/* loss = mean over nodes and output channels of (net(x) - target)^2 (host
 * double); grads (may be null) receives d loss / d params. */
int ld_cnn_grad(ld_cnn* c, const float* x, const float* target, float* grads, double* loss);
/* ld_cnn_grad followed by one Adam step on the parameters. */
int ld_cnn_train_step(ld_cnn* c, const float* x, const float* target, float lr, float beta1, float beta2, float eps,
                      double* loss);

/* Learned-form plumbing (no counterpart in the reference).  All fields are on
 * the context's grid and precision unless stated.
 * ld_spectral_resample: nfields half spectra of an n_src^3 grid -> this grid:
 *   the modes with |kx|, |ky|, kz < min(n_src, N)/2 copied and scaled by
 *   (N / n_src)^3, all others zero (truncation or zero padding).
 * ld_velocity_gradient_spec: gk[3 i + j] = scale * i k_j uk_i, the nine
 *   spectral velocity-gradient components (Nyquist derivatives zero).
 * ld_fields_to_rows: nfields physical fields f[c][node] -> FP32 rows
 *   rows[node][c] * scale (the CNN's input layout).
 * ld_subgrid_compose / ld_subgrid_residual: with ratio r (this grid is the
 *   fine grid, N / r the coarse one) a coarse node's FP32 row of 3 r^3
 *   channels holds the 3 velocity components at its r^3 fine sub-nodes
 *   (channel c r^3 + ((X%r) r + Y%r) r + Z%r):
 *   compose: out = base + scale * net;  residual: net = scale * (fine - base).
 * ld_droplets_frozen_rk4: nsteps RK4 steps of size h of droplet rows (np, 7)
 *   in place in the frozen physical velocity u (3 fields): x' = v,
 *   v' = (I_u(x) - v) / (St tau_eta) - g e_z, I_u the cosine-kernel gather
 *   (particles.py:435-511), St = (m / factor)^(2/3) (particles.py:823), m fixed;
 *   np < 2^29.  Uses the context's droplet gather records as scratch (not
 *   to be called while an ld_advance of the same context is in flight on
 *   another stream). */
int ld_spectral_resample(ld_ctx* ctx, const void* src, int32_t n_src, void* dst, int32_t nfields);
int ld_velocity_gradient_spec(ld_ctx* ctx, const void* uk, void* gk, double scale);
int ld_fields_to_rows(ld_ctx* ctx, const void* f, int32_t nfields, float* rows, float scale);
int ld_subgrid_compose(ld_ctx* ctx, const float* net, int32_t ratio, const void* base, void* out, float scale);
int ld_subgrid_residual(ld_ctx* ctx, const void* fine, const void* base, int32_t ratio, float* net, float scale);
int ld_droplets_frozen_rk4(ld_ctx* ctx, double* rows, int64_t np, const void* u, double h, int32_t nsteps);

#ifdef __cplusplus
}
#endif
#endif /* LD_H */
