/*
 * utcnb.h -- C ABI of the MI355X (gfx950) nbody library.
 *
 * This is the drop-in boundary for the UTC nbody benchapp's hot path: the
 * all-pairs softened-gravity acceleration and the damped kick-drift update.
 * Every entry point takes plain pointers and sizes, returns an int status
 * (0 = success, negative = error; see utcnb_last_error()), and never lets a
 * C++ exception cross the boundary.  Citations are relative to the reference's
 * utc-0.1/ tree.
 *
 * Memory: the caller owns every host buffer and every pointer it passes in;
 * a system handle owns the device buffers it allocates.  A handle is bound to
 * the HIP stream given at creation; calls on one handle must not overlap.
 *
 * Layouts are the reference's: positions T[4N] = (x,y,z,m) per body,
 * velocities T[4N] = (vx,vy,vz,w) per body (benchapps/nbody/seq/nbody.h:125-134).
 */
#ifndef UTCNB_H_
#define UTCNB_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define UTCNB_OK 0
#define UTCNB_EINVAL -1   /* bad argument (size, pointer, enum, value) */
#define UTCNB_EHIP -2     /* a HIP runtime call failed */
#define UTCNB_ENOMEM -3   /* device or host allocation failed */
#define UTCNB_ENODEV -4   /* no usable gfx950 device */
#define UTCNB_ESTATE -5   /* call out of order (e.g. step before set_state) */

/* ---- enums -------------------------------------------------------------- */
/* Arithmetic of the pair interaction and of the integrator state. */
#define UTCNB_PREC_REFERENCE 0 /* the reference's own arithmetic in the state
                                  dtype: sqrt then divide, no fused multiply-add,
                                  j summed in order 0..N-1; bit-identical to
                                  benchapps/nbody/seq/bodysystem.{h:70-96,cc:103-181} */
#define UTCNB_PREC_FP32 1      /* plain FP32: v_rsq_f32 (optionally refined by one
                                  Newton step), FMA, FP32 sums and FP32 state */
#define UTCNB_PREC_MIXED 2     /* FP32 pairs with compensated per-body sums;
                                  FP64 state and integrator */
#define UTCNB_PREC_FP64 3      /* FP64 pairs (correctly rounded sqrt/divide),
                                  compensated FP64 sums, FP64 state */

/* Kick factor applied to the summed field (benchapps/nbody/seq/bodysystem.cc:149,161-163). */
#define UTCNB_KICK_REFERENCE 0 /* v += (a * pos.w) * dt   (the reference: pos.w = m_i) */
#define UTCNB_KICK_PHYSICAL 1  /* v += a * dt */

#define UTCNB_F32 0
#define UTCNB_F64 1

#define UTCNB_HOST 0
#define UTCNB_DEVICE 1

/* config flags */
#define UTCNB_FLAG_NO_NEWTON 1 /* FP32/MIXED: use v_rsq_f32 without the Newton step
                                  (same measured accuracy, ~15% faster) */
#define UTCNB_FLAG_SYM_CHUNKS 2 /* MIXED with a symmetric tile: evaluate each step in
                                   chunks of wave blocks (partial-sum memory for one
                                   chunk's reactions, one FP64 pass more); taken
                                   without the flag when the whole step's partial
                                   sums exceed the device budget */

/* ---- kernel-level entry points (device pointers, reference layout) ------ */
/*
 * One fused step in the reference's arithmetic: forces from pos_old for every
 * body, then v = (v + (a*pos.w)*dt)*damping, x = x + v*dt into pos_new (pos.w
 * copied) and vel (in place).
 * Replaces the launch of _integrateNBodySystem_kernel<T> /
 * _integrateNBodySystemSmall_kernel<T> at
 * benchapps/nbody/utc/sgpu/body_task_sgpu.cu:117-154 (kernels:
 * benchapps/nbody/cuda/bodysystem_kernel.cu:87-138,186-243), with the same
 * argument meaning.  `blocksize` is the workgroup size (64..1024, multiple of
 * 64; 0 = 256).  Unlike the reference, numBodies need not be a multiple of
 * blocksize.  Results are bit-identical to the reference's CPU path; the
 * reference's split-j "Small" variant changes the summation order and is not
 * reproduced.  pos_old and pos_new must not alias.
 */
int utcnb_integrate_f32(const float* pos_old, float* pos_new, float* vel, int numBodies,
                        float deltaTime, float softeningSquared, float damping, int blocksize,
                        void* stream);
int utcnb_integrate_f64(const double* pos_old, double* pos_new, double* vel, int numBodies,
                        double deltaTime, double softeningSquared, double damping,
                        int blocksize, void* stream);

/*
 * The multi-GPU form of the step above, for one GPU's share of the bodies:
 * targets [start, start + count) against all numBodies sources of pos_old
 * (T[4*numBodies]); the slice's new positions go to pos_new[4*count] and its
 * velocities vel[4*count] (slice-local arrays, as the reference's per-GPU
 * buffers) are updated in place.  Replaces the launch of the multi-GPU
 * _integrateNBodySystem_kernel<T> (start_bodyIndex, local_numBodies) at
 * benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:104-131 (kernel:
 * benchapps/nbody/utc-snode-mgpu/mgpu/bodysystem_kernel.cu:88-140), with the
 * same argument meaning; the reference's arithmetic bit for bit.  The caller
 * partitions the bodies (the reference: the first numBodies % G shares take
 * one body more, body_task_mgpu.cu:41-49) and gathers the slices.
 */
int utcnb_integrate_slice_f32(const float* pos_old, float* pos_new, float* vel, int numBodies, int start,
                              int count, float deltaTime, float softeningSquared, float damping, int blocksize,
                              void* stream);
int utcnb_integrate_slice_f64(const double* pos_old, double* pos_new, double* vel, int numBodies, int start,
                              int count, double deltaTime, double softeningSquared, double damping,
                              int blocksize, void* stream);

/* ---- system handle ------------------------------------------------------ */
typedef struct utcnb_system utcnb_system;

typedef struct {
    int32_t n;           /* bodies, 1 .. 2^30 */
    int32_t precision;   /* UTCNB_PREC_* */
    int32_t kick;        /* UTCNB_KICK_* */
    int32_t state_dtype; /* UTCNB_PREC_REFERENCE only: UTCNB_F32 or UTCNB_F64 */
    int32_t flags;       /* UTCNB_FLAG_* */
    int32_t tile;        /* force-kernel tile id 1..utcnb_tile_count() (e.g. from
                            utcnb_autotune); 0 = built-in default.  A
                            small-system tile (name smAtTLwWnCAPk) holds at
                            most CAP bodies (N padded to a multiple of 1024):
                            larger n gives UTCNB_EINVAL */
    double eps2;         /* softening squared, > 0 unless precision is REFERENCE */
    double damping;      /* velocity damping per step (reference default 1.0) */
    double dt;           /* time step */
} utcnb_config;

/*
 * Create a system: allocates the double-buffered device state.  Replaces the
 * GpuData allocations and kernel-shape selection in
 * BodySystemSGPU<T>::runImpl, benchapps/nbody/utc/sgpu/body_task_sgpu.cu:100-127
 * (GpuData: include/gpu/GpuData.inc:69-102).  `stream` is a hipStream_t (NULL =
 * the null stream).
 */
int utcnb_create(const utcnb_config* cfg, void* stream, utcnb_system** out);
/* Free device buffers; replaces the GpuData destructors (include/gpu/GpuData.inc:130-157). */
int utcnb_destroy(utcnb_system* sys);

/*
 * Upload state (pos4 and vel4, n*4 values of `dtype`, converted to the
 * system's state type).  where = UTCNB_HOST for host pointers, UTCNB_DEVICE for
 * device pointers.  Replaces GpuData::initH + sync (H2D) at
 * benchapps/nbody/utc/sgpu/body_task_sgpu.cu:102-111.  vel4 may be NULL
 * (zero velocities, w = 1).  Synchronizes the system's stream when
 * where = UTCNB_HOST, so host buffers may be reused on return.
 */
int utcnb_set_state(utcnb_system* sys, const void* pos4, const void* vel4, int dtype,
                    int where);
/* Download state; replaces GpuData::sync (D2H) + fetch at
 * benchapps/nbody/utc/sgpu/body_task_sgpu.cu:161-170.  Either pointer may be NULL.
 * Synchronizes the system's stream when where = UTCNB_HOST. */
int utcnb_get_state(utcnb_system* sys, void* pos4, void* vel4, int dtype, int where);

/* Advance nsteps steps (asynchronous on the system's stream); replaces the
 * launch loop of benchapps/nbody/utc/sgpu/body_task_sgpu.cu:130-158. */
int utcnb_step(utcnb_system* sys, int nsteps);
/* Wait for the system's stream. */
int utcnb_sync(utcnb_system* sys);

/*
 * One GPU's share of a step of a system sharded over targets (the reference's
 * BodySystemMGPU<T>::runImpl step, benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:98-145):
 * bodies [start, start + count) against the current positions of all n
 * bodies; the slice's new positions go to the next position buffer, its
 * velocities (and, for MIXED, its FP64 state) are updated in place.  The next
 * buffer is not made current: the caller fills the other slices of it (its own
 * further utcnb_step_slice calls, or the peers' slices received into
 * utcnb_position_buffer(sys, 1, ...) by an all-gather), then calls
 * utcnb_commit_step.  FP32 and MIXED run a one-pass tile (utcnb_slice_tile):
 * the configured tile when it is one, else a default by count; REFERENCE and
 * FP64 run their only kernels.  Per body the arithmetic does not depend on the
 * slicing: slices covering [0, n) then a commit give bitwise the same state as
 * utcnb_step(sys, 1) with that one-pass tile.  Asynchronous on the handle's
 * stream; utcnb_last_kernel_ms times the slice.
 */
int utcnb_step_slice(utcnb_system* sys, int start, int count);
/* Make the next position buffer current (after the slices of a step). */
int utcnb_commit_step(utcnb_system* sys);
/*
 * Device address of the current (which = 0) or next (which = 1) position
 * buffer: n records of 4 values (x, y, z, m) of *dtype (UTCNB_F32 for FP32,
 * MIXED and REFERENCE/F32 -- the FP32 source snapshot for MIXED -- UTCNB_F64
 * otherwise), the array an all-gather of slices writes.  No reference
 * counterpart (the reference gathers slices through host memory,
 * body_task_mgpu.cu:109,134-140).
 */
int utcnb_position_buffer(utcnb_system* sys, int which, void** pos4, int* dtype);
/*
 * The same share of a step for a MIXED system with a symmetric tile, which
 * evaluates each unordered pair of bodies once: the shares are wave blocks
 * (64*K bodies) rather than single bodies, and a pair of bodies owned by two
 * shares is evaluated by one of them, which computes the reaction on the
 * other's body.  A step of share `share` of `shares`:
 *   1. utcnb_share_pairs: the pair work of the share's wave blocks; the
 *      reactions it computes are summed per body into react3 (device FP64
 *      [3n], written in full; the caller's buffer);
 *   2. the caller sums react3 over the shares for the share's own bodies
 *      (a reduce-scatter of the shares' react3 arrays in a multi-GPU run;
 *      on one device, a plain sum);
 *   3. utcnb_share_finish: the share's bodies [start, start + count)
 *      (utcnb_share_range) add their target-side partials to react_sum3
 *      (device FP64 [3*count]; may be NULL when count is 0, a share of
 *      padding bodies only) and are integrated; their new positions go to
 *      the next position buffer;
 *   4. the caller gathers the other shares' positions into that buffer and
 *      calls utcnb_commit_step.
 * Every share's partial sums are those of utcnb_step with the same tile; only
 * the FP64 order of the final per-body sums differs.  No reference
 * counterpart (the reference's multi-GPU kernel evaluates every ordered pair,
 * benchapps/nbody/utc-snode-mgpu/mgpu/bodysystem_kernel.cu:88-140).
 * 1 <= shares <= the number of wave blocks; asynchronous on the handle's stream.
 */
int utcnb_share_range(const utcnb_system* sys, int shares, int share, int* start, int* count);
int utcnb_share_pairs(utcnb_system* sys, int shares, int share, double* react3);
int utcnb_share_finish(utcnb_system* sys, int shares, int share, const double* react_sum3);
/* Field of bodies [start, start + count) as utcnb_step_slice computes it:
 * acc3 host FP64 [3 * count].  Synchronizes the stream. */
int utcnb_accel_slice(utcnb_system* sys, int start, int count, double* acc3);
/* Tile id utcnb_step_slice / utcnb_accel_slice run for `count` targets (0 for
 * REFERENCE and FP64).  No reference counterpart. */
int utcnb_slice_tile(const utcnb_system* sys, int count, int* tile);

/* Acceleration field of the current state, as the system's force kernel
 * computes it (FP32 hi+lo for MIXED), written as FP64 acc3[3n] (host). */
int utcnb_accel(utcnb_system* sys, double* acc3);
/*
 * The field of utcnb_accel for n bodies given on the device, with no handle
 * and no device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  pos4: device FP32 [4n] (x,y,z,m);
 * acc3: device FP64 [3n]; asynchronous on `stream` (a hipStream_t, NULL =
 * the null stream).  precision FP32 (plain sums) or MIXED (compensated), tile
 * and flags (UTCNB_FLAG_NO_NEWTON) as in utcnb_config; the result is bitwise
 * the utcnb_accel field of a handle of that precision, tile and flags whose
 * FP32 source positions equal pos4.  Work memory (padded sources, partial
 * sums) comes from a 64 MiB device array that the library holds on each
 * device it runs on until the process ends; sizes whose work memory exceeds
 * it are refused with UTCNB_EINVAL (with the default tile, n up to about
 * 32768 on the symmetric tiles and about 4M on the one-pass ones).  Calls on
 * one device share that array: calls that overlap on one device race.
 */
int utcnb_accel_noalloc(const float* pos4, int n, double eps2, int precision, int tile, int flags,
                        double* acc3, void* stream);
/* FP64 diagnostics of the current state: kinetic 1/2 sum m|v|^2, potential
 * -sum_{i<j} m_i m_j / sqrt(r^2+eps2) and momentum sum m v (host outputs). */
int utcnb_diagnostics(utcnb_system* sys, double* kinetic, double* potential,
                      double* momentum3);

/*
 * Forward `nsteps` steps from the current state, then the adjoint for
 * L = (1/n) sum_i |x_i(T) - t_i|^2 back to that state.  FP64 checkpoints of
 * (x, v) are kept every `ckpt_every` steps and the steps in between are
 * recomputed during the reverse sweep, except those of the last segment,
 * whose positions the forward sweep leaves in the segment buffer (the same
 * bits a recompute gives).  With more than one segment the recomputes run on
 * an internal stream, one segment ahead of the VJP steps on the handle's
 * stream (the reference kick keeps them on the handle's stream); the call
 * returns after both.  Precision MIXED (FP32 pair VJP) or
 * FP64.  target3: host FP64 [3n].  Outputs (host, FP64): *loss, dx0[3n],
 * dv0[3n], dm[n].  The system ends in the final (step nsteps) state.
 */
int utcnb_adjoint(utcnb_system* sys, int nsteps, int ckpt_every, const double* target3,
                  double* loss, double* dx0, double* dv0, double* dm);

/* Total time (ms) of the force kernels launched by the last utcnb_step /
 * utcnb_adjoint call on this handle, measured with HIP events on its stream;
 * *launches receives their count.  Synchronizes the stream. */
int utcnb_last_kernel_ms(utcnb_system* sys, double* ms, int* launches);

/* Chunks of wave blocks a MIXED handle's symmetric steps are evaluated in
 * (UTCNB_FLAG_SYM_CHUNKS, or a system whose partial sums exceed the budget in
 * one launch); 0 for a single launch or no symmetric tile.  No reference
 * counterpart. */
int utcnb_sym_chunks(const utcnb_system* sys, int* chunks);

/* Force-kernel tile id (1..utcnb_tile_count()) the handle runs: the configured
 * one, or the built-in default for its size when utcnb_config.tile was 0; 0
 * for the REFERENCE and FP64 precisions, whose kernels have no tiles.
 * No reference counterpart (the reference's block size is a launch argument). */
int utcnb_get_tile(const utcnb_system* sys, int* tile);

/* ---- batched systems ---------------------------------------------------- */
typedef struct utcnb_batch utcnb_batch;

/*
 * Many independent systems of uneven size in one device allocation.
 * sizes[nsys]; every system shares cfg's precision (MIXED only), kick, eps2,
 * damping and dt (cfg->n is ignored).  The reference's analog is
 * benchapps/nbody/utc/nbody_mtask.cc:185-231 (one task and stream per system).
 */
int utcnb_batch_create(int nsys, const int32_t* sizes, const utcnb_config* cfg, void* stream,
                       utcnb_batch** out);
int utcnb_batch_destroy(utcnb_batch* b);
/* Concatenated FP64 host state, systems in order: pos4/vel4 [4 * sum(sizes)]. */
int utcnb_batch_set_state(utcnb_batch* b, const double* pos4, const double* vel4);
int utcnb_batch_get_state(utcnb_batch* b, double* pos4, double* vel4);
/*
 * mode 0: one captured HIP graph of persistent step kernels over all systems
 *         (captured on first use for a given nsteps/ckpt_every/adjoint, then
 *         replayed; every run starts from the FP64 state);
 * mode 1: per-system launches of the single-system kernels (the baseline);
 * mode 2: the persistent kernels of mode 0 launched eagerly (no graph).
 * cfg->tile of the batch selects the batched kernels: 1..3 = one-pass with 4,
 * 8 or 16 waves per (system, 64-target tile) work item on persistent
 * workgroups; 4 = symmetric (each unordered pair once) with (system, wave
 * block, split) items, longest first, taken by persistent waves through an
 * atomic work counter; 5 = symmetric, one wave per item;
 * 0 = time all five on the first mode 0/2 run and keep the fastest.  With
 * tile 0, 4 or 5 every system is padded to a multiple of 256 bodies, else 64.
 * The batched adjoint supports the physical kick only.
 * When target3 is non-NULL the adjoint of the same loss as utcnb_adjoint runs
 * after the forward steps (checkpoint every ckpt_every) and losses[nsys],
 * dx0/dv0 [3 * sum(sizes)], dm [sum(sizes)] receive per-system results.
 * *ms receives the wall time of the call measured with HIP events.
 */
int utcnb_batch_run(utcnb_batch* b, int nsteps, int mode, int ckpt_every,
                    const double* target3, double* losses, double* dx0, double* dv0,
                    double* dm, double* ms);
/* The batch tile (1..5, see above) the batch runs: the configured one, the
 * tuned one after the first mode 0/2 run, or 0 before tuning.  No reference
 * counterpart. */
int utcnb_batch_get_tile(const utcnb_batch* b, int* tile);

/* ---- tile autotuner ----------------------------------------------------- */
/*
 * Time the force-kernel tiles for an n-body system in `precision` on the
 * current device and return the fastest tile id (usable as utcnb_config.tile).
 * Results are cached in memory and, when cache_path is non-NULL, in a text file
 * keyed by device and problem class.
 */
int utcnb_autotune(int n, int precision, const char* cache_path, int* best_tile,
                   double* best_ms);
/* Number of tiles and a description of tile id t (static string). */
int utcnb_tile_count(void);
const char* utcnb_tile_name(int tile);

/* ---- input generators (host) -------------------------------------------- */
/* xoshiro256** seeded by splitmix64 (published algorithms of Blackman & Vigna). */
void utcnb_rng_seed(uint64_t seed, uint64_t state[4]);
uint64_t utcnb_rng_next(uint64_t state[4]);
uint64_t utcnb_splitmix64(uint64_t* x);
/*
 * Fill pos4/vel4 (FP64 [4n]) and, for TARGETS, out3 [3n]:
 *   UTCNB_GEN_CUBE      x ~ U[-1,1]^3, v ~ N(0, p0^2), m = 1/n, w = 1
 *   UTCNB_GEN_CUBE_LOGM as CUBE with m ~ exp(U[ln p1, ln p2]) / n
 *   UTCNB_GEN_PLUMMER   Plummer sphere, a = 1, total mass 1, equal masses,
 *                       Aarseth-Henon-Wielen (1974) sampler, mass cut 0.999,
 *                       centre of mass and mean velocity removed
 *   UTCNB_GEN_TARGETS   out3 ~ N(0, p0^2)
 */
#define UTCNB_GEN_CUBE 0
#define UTCNB_GEN_CUBE_LOGM 1
#define UTCNB_GEN_PLUMMER 2
#define UTCNB_GEN_TARGETS 3
int utcnb_generate(int kind, int n, uint64_t seed, double p0, double p1, double p2,
                   double* pos4, double* vel4, double* out3);

/* ---- misc --------------------------------------------------------------- */
/* Last error message of the calling thread ("" if none). */
const char* utcnb_last_error(void);
/* Library version string. */
const char* utcnb_version(void);
/* Device name and compute-unit count of the current device (0 on success). */
int utcnb_device_info(char* name, int name_len, int* cu_count, int* clock_khz);

#ifdef __cplusplus
}
#endif
#endif /* UTCNB_H_ */
