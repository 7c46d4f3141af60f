/*
 * mog.h -- C ABI of the MI355X (gfx950) Mixture-of-Grids force step.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses
 * (for example torch.Tensor.data_ptr() of a CUDA/HIP tensor); `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  Every call is
 * asynchronous on `stream` unless stated otherwise.
 *
 * The reference (tbmo/mog-solver) has no FFI layer: its solver interface is
 * the `Simulation` class of main.py, which drives GLSL compute shaders through
 * moderngl.  Each entry point below names the reference code it replaces.
 *
 * Return codes: 0 = MOG_OK, negative = error (see mog_status); the message of
 * the last error of a context is returned by mog_last_error(ctx), and of a
 * failed mog_create by mog_last_error(NULL).
 *
 * Threading: a context is not re-entrant; one context per device and stream.
 * Ownership: the caller owns particle buffers; the context owns density
 * grids, force grids, transform copies and tables.
 */
#ifndef MOG_H
#define MOG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOG_ABI_VERSION 2

typedef enum mog_status {
  MOG_OK = 0,
  MOG_EINVAL = -1,        /* invalid argument (size, pointer, alignment, config) */
  MOG_EHIP = -2,          /* HIP runtime error */
  MOG_ENOTRANSFORMS = -3, /* mog_set_transforms was not called */
  MOG_EUNSUPPORTED = -4,  /* valid but not implemented option */
  MOG_ENOMEM = -5         /* device allocation failed */
} mog_status;

enum { MOG_COMPAT_REFERENCE = 0, MOG_COMPAT_PHYSICAL = 1 };
enum { MOG_PRECISION_FP32 = 0, MOG_PRECISION_BF16X3 = 1, MOG_PRECISION_FP16 = 2 };
enum { MOG_SOLVER_FFT = 0, MOG_SOLVER_CONVOLUTION = 1 };

/* Solver configuration.  Keys mirror config.yaml as read by
 * main.py:115-126 and main.py:662-667 (n_grids, grid_size, world_size, G,
 * dt, damping, solver) plus the deltaTime multiplier `timescale`
 * (main.py:111, 662) and the build's own options (compat, precision). */
typedef struct mog_config {
  int32_t n_grids;    /* k in [1, 65535] */
  int32_t grid_size;  /* g, power of two in [2, 64] */
  double world_size;  /* W > 0, periodic cube side */
  double G;           /* gravitational constant of greens.glsl */
  double dt;          /* time step */
  double damping;     /* velocity factor per step */
  double timescale;   /* deltaTime = dt * timescale */
  int32_t compat;     /* MOG_COMPAT_*: reference keeps the x<-z force aliasing of
                         main.py:638-654 at g in {4,16,64}; physical does not */
  int32_t precision;  /* MOG_PRECISION_* of the DFT contractions */
  int32_t solver;     /* MOG_SOLVER_*; the reference runs FFT for both */
  int32_t reserved;
} mog_config;

/* Launch and algorithm choices of a context (defaults from mog_create;
 * results do not depend on them, only speed does).  The reference has no
 * counterpart: its dispatch sizes are fixed (main.py:585, 596, 607, 668). */
typedef struct mog_tuning {
  int32_t deposit_threads;      /* threads per deposit workgroup: 64..1024, multiple of 64 */
  int32_t deposit_grids_per_wg; /* grids privatised per workgroup: 1, 2, 4 or 8 */
  int32_t deposit_wg_target;    /* approximate number of deposit workgroups, >= 1 */
  int32_t resort_every;         /* steps between rebuilds of the Morton visiting order; 0 = off */
  int32_t sorted_deposit;       /* 1: deposit positions in visiting order, wave-aggregated adds */
  int32_t cell_cache;           /* 1: the sorted deposit stores cells for the gather (g <= 32) */
  int32_t spectral_mfma;        /* 1: matrix-core DFT for FP32 at g >= 16 (other precisions
                                   always use it) */
  int32_t deposit_prepass;      /* LDS deposit reads rows prepared by a prepass (centred
                                   coordinates, milli-mass): 0 = when n*k >= 2^26 (default),
                                   1 = always, -1 = never; results are identical */
  int32_t reserved[8];
} mog_tuning;

typedef struct mog_ctx mog_ctx;

/* ABI version of the loaded library (MOG_ABI_VERSION). */
int mog_abi_version(void);

/* Validate `cfg`, allocate the context's device buffers on the current HIP
 * device and build the FP64-derived tables.  Replaces Simulation.__init__'s
 * buffer/texture/shader set-up (main.py:108-126, 191-248). */
int mog_create(const mog_config* cfg, mog_ctx** out);

/* Release a context (NULL is ignored). */
void mog_destroy(mog_ctx* ctx);

/* Copy per-grid transforms from HOST memory (synchronous): rot is (k,12)
 * float32 with rot[i*12 + 4*j + m] = M_i[j][m], off is (k,4) float32.
 * Replaces the offset/rotation SSBO uploads of main.py:196-197 fed by
 * compute_transforms (main.py:157-172). */
int mog_set_transforms(mog_ctx* ctx, const float* rot_k12, const float* off_k4);

/* One full force step on device particle state, in place:
 * pos (n,4) float32 = (x, y, z, mass), vel (n,4) float32 = (vx, vy, vz, 0),
 * both 16-byte aligned.  Replaces Simulation.step / step_fft
 * (main.py:592-672): clear, deposit, spectral solve, gather + update. */
int mog_step(mog_ctx* ctx, float* pos_n4, float* vel_n4, int64_t n, void* stream);

/* `nsteps` consecutive steps; with use_graph != 0 one step is captured in a
 * hipGraph on first use for these (pos, vel, n, stream) and replayed.  From
 * 4 steps on, the run works on a copy of the state in the gather's visiting
 * order (the context's own buffers): vel (and, in graph replays that feed
 * the deposit prepared rows, pos) is written back at each rebuild of the
 * order and at the end of the call's work on the stream, with the same bits
 * as nsteps mog_step calls; the arrays are current once that work is done. */
int mog_run(mog_ctx* ctx, float* pos_n4, float* vel_n4, int64_t n, int32_t nsteps,
            int32_t use_graph, void* stream);

/* Phase entry points (tests and callers that stage the step themselves). */

/* Zero `grids_out` (k*g^3 int64, [grid][z][y][x]) and deposit the particles'
 * milli-mass uint(mass*1000) by NGP.  Replaces clear_mass (main.py:558-562)
 * + mass.glsl dispatch (main.py:599-608). */
int mog_deposit(mog_ctx* ctx, const float* pos_n4, int64_t n, int64_t* grids_out, void* stream);

/* Force grids (k*g^3 float4 = Fx, Fy, Fz, |F|) from density grids.  Replaces
 * complex.glsl, run_fft_3d_batched, fourier.glsl, greens.glsl, the three
 * inverse transforms and pack.glsl (main.py:610-655). */
int mog_solve(mog_ctx* ctx, const int64_t* grids, float* force_out, void* stream);

/* Averaged world-frame force avgForce (n,4; w = 0) at the particles, from
 * force grids.  The gather half of update.glsl:65-81. */
int mog_gather(mog_ctx* ctx, const float* pos_n4, int64_t n, const float* force,
               float* avg_force_n4, void* stream);

/* Kick, damp, drift, wrap from a given avgForce.  The update half of
 * update.glsl:83-89. */
int mog_update(mog_ctx* ctx, float* pos_n4, float* vel_n4, const float* avg_force_n4, int64_t n,
               void* stream);

/* mog_step without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  mog_step
 * allocates the gather's Morton visiting order (keys, indices, radix-sort
 * temporaries) on its first call for a particle count; this variant keeps
 * it in a device-resident array of the library instead, 20 MiB per GPU
 * (held until the process ends), shared by every call on that GPU and
 * keeping the last call's contents.  Limit: n <= 1048576 (2^20) particles,
 * larger n returns MOG_EINVAL before any launch.  Results are bit-identical
 * to mog_step (the sorted-deposit tuning is not used here; it does not
 * change results).  Calls that overlap on one GPU (other threads or
 * streams, any context) race on that array. */
int mog_step_noalloc(mog_ctx* ctx, float* pos_n4, float* vel_n4, int64_t n, void* stream);

/* mog_step with HIP events recorded between its kernels on `stream`; waits
 * for the stream and writes the durations in milliseconds of the phases
 * {grid clear, deposit, spectral solve, gather+update} to phase_ms[0..3].
 * The clear phase holds no kernel (only the sorted-deposit tuning's
 * permutation): the grids a step deposits into were cleared by the
 * previous step's gather+update.  The reference has no timing facility;
 * this exists for the benchmark. */
int mog_step_profiled(mog_ctx* ctx, float* pos_n4, float* vel_n4, int64_t n, float* phase_ms,
                      void* stream);

/* Change deltaTime = dt * timescale and the damping factor between steps
 * (main.py:662, 667 read them every step; the reference's [ ] \ keys change
 * timescale at run time, main.py:739-750).  Drops a captured mog_run graph. */
int mog_set_integrator(mog_ctx* ctx, double dt, double damping, double timescale);

/* Perlin-fBm rejection-sampled initial conditions on the device:
 * shaders/init_particles.glsl driven by Simulation.init_particles_gpu
 * (main.py:283-310): positions in a sphere (spawn_shape 1) or cube (0) of the
 * box shrunk by spawn_buffer * W, density pow(clamp(fbm4 + 0.5, 0, 1),
 * density_contrast) at noise frequency noise_scale / W, masses base_mass,
 * velocities uniform in [-W/2, W/2).  No context needed. */
int mog_init_perlin(float* pos_n4, float* vel_n4, int64_t n, double world_size,
                    double noise_scale, double density_contrast, uint32_t seed,
                    double base_mass, double spawn_buffer, int32_t spawn_shape, void* stream);

/* FP64 diagnostics of a particle state (no counterpart in the reference,
 * which prints only its configuration, main.py:147-155).  Writes
 * out8[0..7] (DEVICE pointer, 8 doubles, async on `stream`) =
 *   {sum m vx, sum m vy, sum m vz, 1/2 sum m |v|^2, U, sum m, 0, 0}
 * with m = pos.w and U the mean over the k grids of the potential energy
 * 1/2 sum_c rho(c) phi(c) of the NGP density of `pos`: rho the milli-mass
 * per cell / 1000 (complex.glsl:23) and phi_hat = -2 pi G rho_hat / k^2
 * (greens.glsl:40-47, float32 k2 > 1e-4 mask, DC zeroed), the potential
 * whose gradient the force grids hold (the step accelerates by -grad phi,
 * update.glsl:83).  U is evaluated as sum_k T(k) |m_hat(k)|^2 by Parseval,
 * every operation FP64, sums in a fixed order.  Uses the context's spare
 * density grids (left all zero) and allocates its tables on the first call. */
int mog_diagnostics(mog_ctx* ctx, const float* pos_n4, const float* vel_n4, int64_t n, double* out8,
                    void* stream);

/* ---- Learned closure (no counterpart in the reference) ----------------
 * Features of the ensemble of `ctx` (typically few grids, e.g. k = 24) at
 * the particles: runs the context's clear, deposit and spectral solve, a
 * second solve with the Green's function divided by the NGP window
 * sinc(k_x h/2) sinc(k_y h/2) sinc(k_z h/2), and writes
 * feat (n,12) = (m/S, s/S, c/S, log(1 + rho/rho0), S, 0) and
 * mean (n,4) = (c, 0), with m the mean and s the across-grid standard
 * deviation of the world-frame NGP grid forces, c the mean trilinear
 * interpolation of the deconvolved world-frame force fields, rho the mean
 * mass of the particle's cells and S = |c| + |s| a per-particle force
 * scale. */
int mog_closure_features(mog_ctx* ctx, const float* pos_n4, int64_t n, double rho0,
                         float* feat_n12, float* mean_n4, void* stream);

/* Number of fp32 parameters of the closure MLP (10 -> 64 -> 64 -> 64 -> 3,
 * SiLU; W1 b1 W2 b2 W3 b3 W4 b4 with W[in][out] row-major). */
int mog_mlp_param_count(void);

/* force (n,4) = mean + S * MLP(x) on the matrix cores (bf16 != 0: BF16
 * operands, else FP32; FP32 accumulation); x = feat[0..9], S = feat[10]. */
int mog_mlp_forward(const float* params, const float* feat_n12, const float* mean_n4, int64_t n,
                    float* force_n4, int32_t bf16, void* stream);

/* Fused forward + backward over n particles: adds dLoss/dparams to grad
 * and the loss to *loss (device pointers; the caller clears them), with
 * loss = wnorm * sum_p S_p^2 |MLP(x_p) - (target_p - mean_p)/S_p|^2. */
int mog_mlp_train_grad(const float* params, const float* feat_n12, const float* mean_n4,
                       const float* target_n4, int64_t n, double wnorm, float* grad, float* loss,
                       int32_t bf16, void* stream);

/* FP64 direct summation audit: at the particles targets[0..nt) of pos,
 * grad_out (nt,3) = G_eff * sum_j m_j d / (|d|^2 + softening^2)^1.5 with d the
 * minimum-image separation in the periodic box of side world_size (the
 * gradient the MOG avgForce approximates, G_eff = G h^3 / 2). */
int mog_direct_sum(const float* pos_n4, int64_t n, const int32_t* targets, int32_t nt,
                   double world_size, double G_eff, double softening, double* grad_out,
                   void* stream);

/* One bias-corrected Adam update of the parameters (step >= 1). */
int mog_mlp_adam(float* params, const float* grad, float* m1, float* m2, double lr, double beta1,
                 double beta2, double eps, int32_t step, void* stream);

/* Current tuning of `ctx` (host struct, synchronous). */
int mog_get_tuning(mog_ctx* ctx, mog_tuning* out);

/* Replace the tuning of `ctx`; invalid fields are rejected with MOG_EINVAL
 * and nothing changes.  Drops a captured mog_run graph. */
int mog_set_tuning(mog_ctx* ctx, const mog_tuning* t);

/* Device pointers of the context's own density grids (int64) and force
 * grids (float4) as left by the last mog_step; NULL on error.  The density
 * pointer alternates between two buffers from one step to the next, so it
 * is valid until the next step.  After a step that returned an error their
 * contents are unspecified until the next successful step. */
const int64_t* mog_grids_ptr(mog_ctx* ctx);
const float* mog_force_ptr(mog_ctx* ctx);

/* ---- Batched ensemble (no counterpart in the reference) -----------------
 * n_sims independent simulations stepped together: one persistent launch
 * per phase (deposit, three solve passes, gather+update) over all of them,
 * optionally replayed as one hipGraph.  Simulation s has its own config
 * (grid_size <= 32, fp32; any n_grids, compat, solver, integrator) and
 * particle count n[s]; its particles are rows [off_s, off_s + n[s]) of the
 * packed pos/vel arrays, off_s = n[0] + ... + n[s-1].  Stepping the batch
 * gives the same state, bit for bit, as stepping each simulation alone with
 * the FP32 vector spectral solve.  Errors are reported through
 * mog_last_error(NULL). */
typedef struct mog_batch mog_batch;

int mog_batch_create(const mog_config* cfgs, const int64_t* n, int32_t n_sims, mog_batch** out);
void mog_batch_destroy(mog_batch* batch);

/* Transforms of simulation `sim` (host arrays, copied), as mog_set_transforms. */
int mog_batch_set_transforms(mog_batch* batch, int32_t sim, const float* rot_k12,
                             const float* off_k4);

/* One step of every simulation (async on `stream`). */
int mog_batch_step(mog_batch* batch, float* pos_n4, float* vel_n4, void* stream);

/* nsteps steps; use_graph != 0 captures the step once per (pos, vel) and
 * replays it. */
int mog_batch_run(mog_batch* batch, float* pos_n4, float* vel_n4, int32_t nsteps,
                  int32_t use_graph, void* stream);

/* Deposit task shape: particles per task (256 .. 2^24) and threads per
 * workgroup (256, 512, 1024).  Rebuilds the task list, drops the graph. */
int mog_batch_set_tuning(mog_batch* batch, int64_t deposit_chunk, int32_t deposit_threads);
int mog_batch_get_tuning(const mog_batch* batch, int64_t* deposit_chunk, int32_t* deposit_threads,
                         int32_t* n_tasks);

/* ---- 2D path (the DIM == 2 branches of update.glsl:19-34, pack.glsl:8;
 * never compiled by the reference, main.py:237) -------------------------
 * k planar grids of g x g cells over [0, W)^2.  rot_k12 / off_k4 use the 3D
 * packed layout; the 2x2 block rot[4j+m] (j, m < 2) is the planar rotation
 * (mat2(col0.xy, col1.xy)) and off[0..1] the offset.  Particles are (n,4)
 * float32 = (x, y, unused, m).  Density is int64 milli-mass per cell, the
 * Green's function -2 pi G / k^2 (DC removed), the force i k phi; the
 * average world force is written as (Fx, Fy, 0, 0).  Errors are reported
 * through mog_last_error(NULL). */
typedef struct mog2d_ctx mog2d_ctx;

int mog2d_create(int32_t n_grids, int32_t grid_size, double world_size, double G,
                 const float* rot_k12, const float* off_k4, mog2d_ctx** out);
void mog2d_destroy(mog2d_ctx* ctx);

/* Average force at the particles (async on `stream`). */
int mog2d_forces(mog2d_ctx* ctx, const float* pos_n4, int64_t n, float* avg_n4, void* stream);

/* Clear + deposit only: the k g^2 int64 milli-mass grids ([grid][y][x]) into
 * grids_out (device).  No counterpart in the reference (whose 2D branch has
 * no deposit); the deposit half of mog2d_forces, for tests and callers that
 * want the density. */
int mog2d_deposit(mog2d_ctx* ctx, const float* pos_n4, int64_t n, int64_t* grids_out, void* stream);

/* One step: force, then vel -= F dt; vel *= damping; pos += vel dt; pos mod W
 * (update.glsl:83-86), z untouched when vel.z is 0. */
int mog2d_step(mog2d_ctx* ctx, float* pos_n4, float* vel_n4, int64_t n, double dt, double damping,
               void* stream);

/* Message of the last error of `ctx` (or of the last failed mog_create when
 * ctx is NULL).  Never NULL. */
const char* mog_last_error(const mog_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* MOG_H */
