/*
 * nbody_oracle.c -- CPU oracle for the UTC nbody hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, loaded by the
 * tests, the smoke entry point and the CPU-baseline leg of the benchmark.
 * The product path never links it.
 *
 * Build with -O2 -ffp-contract=off and no -mfma: the reference's CPU path is
 * plain IEEE single/double arithmetic in source order, and bit parity with its
 * golden frame file depends on no multiply-add being fused.
 *
 * Citations are relative to the reference's utc-0.1/ tree.
 */
#define _GNU_SOURCE
#include "nbody_oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Pair interaction, benchapps/nbody/seq/bodysystem.h:70-96.
 * r = pj - pi; d2 = (rx*rx + ry*ry) + rz*rz; d2 += eps2;
 * inv = 1.0 / sqrt_T(d2)   -- for T=float: sqrtf, then a DOUBLE division
 *                              (the literal 1.0 is double), rounded to float;
 * inv3 = (inv*inv)*inv; s = mj*inv3; acc += r*s.                            */
static inline void pair_f32(float acc[3], const float* pi, const float* pj, float eps2)
{
    float r0 = pj[0] - pi[0];
    float r1 = pj[1] - pi[1];
    float r2 = pj[2] - pi[2];
    float d2 = r0 * r0 + r1 * r1 + r2 * r2;
    d2 += eps2;
    float inv = (float)(1.0 / (double)sqrtf(d2));
    float inv3 = inv * inv * inv;
    float s = pj[3] * inv3;
    acc[0] += r0 * s;
    acc[1] += r1 * s;
    acc[2] += r2 * s;
}

static inline void pair_f64(double acc[3], const double* pi, const double* pj, double eps2)
{
    double r0 = pj[0] - pi[0];
    double r1 = pj[1] - pi[1];
    double r2 = pj[2] - pi[2];
    double d2 = r0 * r0 + r1 * r1 + r2 * r2;
    d2 += eps2;
    double inv = 1.0 / sqrt(d2);
    double inv3 = inv * inv * inv;
    double s = pj[3] * inv3;
    acc[0] += r0 * s;
    acc[1] += r1 * s;
    acc[2] += r2 * s;
}

/* BodySystem<T>::_computeNBodyGravitation, seq/bodysystem.cc:103-133.
 * The reference unrolls j by 4 with no remainder loop (it requires n%4==0);
 * the summation order is plain j = 0..n-1, which is what is restated here. */
static void grav_rows_f32(int n, const float* pos4, float eps2, int i0, int i1, float* force3)
{
    for (int i = i0; i < i1; ++i) {
        float acc[3] = {0.f, 0.f, 0.f};
        const float* pi = pos4 + 4 * (size_t)i;
        for (int j = 0; j < n; ++j)
            pair_f32(acc, pi, pos4 + 4 * (size_t)j, eps2);
        force3[3 * (size_t)i + 0] = acc[0];
        force3[3 * (size_t)i + 1] = acc[1];
        force3[3 * (size_t)i + 2] = acc[2];
    }
}

void orc_gravitation_f32(int n, const float* pos4, float eps2, float* force3)
{
    grav_rows_f32(n, pos4, eps2, 0, n, force3);
}

static void grav_rows_f64(int n, const double* pos4, double eps2, int i0, int i1, double* force3)
{
    for (int i = i0; i < i1; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        const double* pi = pos4 + 4 * (size_t)i;
        for (int j = 0; j < n; ++j)
            pair_f64(acc, pi, pos4 + 4 * (size_t)j, eps2);
        force3[3 * (size_t)i + 0] = acc[0];
        force3[3 * (size_t)i + 1] = acc[1];
        force3[3 * (size_t)i + 2] = acc[2];
    }
}

void orc_gravitation_f64(int n, const double* pos4, double eps2, double* force3)
{
    grav_rows_f64(n, pos4, eps2, 0, n, force3);
}

/* Threaded rows.  Body partition follows
 * benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:41-49: with T workers
 * and c rows, the first c%T workers take c/T+1 rows, the rest c/T.  Worker t
 * is pinned to the t-th CPU of the process's affinity set (round robin). */
typedef struct {
    int n, i0, i1;
    const float* pos4;
    float eps2;
    float* force3;
    int cpu; /* CPU to pin this worker to, -1: leave it to the scheduler */
} rows_job;

static void* rows_worker(void* p)
{
    rows_job* jb = (rows_job*)p;
    if (jb->cpu >= 0) {
        cpu_set_t set;
        CPU_ZERO(&set);
        CPU_SET(jb->cpu, &set);
        (void)pthread_setaffinity_np(pthread_self(), sizeof(set), &set); /* best effort */
    }
    grav_rows_f32(jb->n, jb->pos4, jb->eps2, jb->i0, jb->i1, jb->force3);
    return NULL;
}

/* The k-th CPU this process may run on (round robin), or -1. */
static int allowed_cpu(int k)
{
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) != 0)
        return -1;
    int count = CPU_COUNT(&set);
    if (count <= 0)
        return -1;
    k %= count;
    for (int c = 0; c < CPU_SETSIZE; ++c)
        if (CPU_ISSET(c, &set) && k-- == 0)
            return c;
    return -1;
}

void orc_gravitation_rows_f32(int n, const float* pos4, float eps2, int i0, int i1,
                              float* force3, int nthreads)
{
    int cnt = i1 - i0;
    if (cnt <= 0)
        return;
    if (nthreads <= 1 || cnt < nthreads) {
        grav_rows_f32(n, pos4, eps2, i0, i1, force3);
        return;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    rows_job* jobs = (rows_job*)malloc(sizeof(rows_job) * (size_t)nthreads);
    int per = cnt / nthreads, extra = cnt % nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int local = t < extra ? per + 1 : per;
        int start = t < extra ? t * (per + 1) : t * per + extra;
        jobs[t].n = n;
        jobs[t].i0 = i0 + start;
        jobs[t].i1 = i0 + start + local;
        jobs[t].pos4 = pos4;
        jobs[t].eps2 = eps2;
        jobs[t].force3 = force3;
        jobs[t].cpu = allowed_cpu(t);
        pthread_create(&th[t], NULL, rows_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t)
        pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* BodySystem<T>::_integrateNBodySystem, seq/bodysystem.cc:135-181.
 * The kick factor is pos.w (named invMass in the reference, holds m_i);
 * v += (F*w)*dt; v *= damping; x += v*dt.  pos.w and vel.w are untouched. */
void orc_integrate_f32(int n, float* pos4, float* vel4, const float* force3, float dt,
                       float damping)
{
    for (int i = 0; i < n; ++i) {
        float* p = pos4 + 4 * (size_t)i;
        float* v = vel4 + 4 * (size_t)i;
        const float* f = force3 + 3 * (size_t)i;
        float w = p[3];
        for (int k = 0; k < 3; ++k) {
            float vk = v[k];
            vk += (f[k] * w) * dt;
            vk *= damping;
            p[k] += vk * dt;
            v[k] = vk;
        }
    }
}

void orc_integrate_f64(int n, double* pos4, double* vel4, const double* force3, double dt,
                       double damping)
{
    for (int i = 0; i < n; ++i) {
        double* p = pos4 + 4 * (size_t)i;
        double* v = vel4 + 4 * (size_t)i;
        const double* f = force3 + 3 * (size_t)i;
        double w = p[3];
        for (int k = 0; k < 3; ++k) {
            double vk = v[k];
            vk += (f[k] * w) * dt;
            vk *= damping;
            p[k] += vk * dt;
            v[k] = vk;
        }
    }
}

/* BodySystem<T>::update, seq/bodysystem.cc:64-73 (Jacobi: all forces from the
 * old positions, then the integrator). */
int orc_run_f32(int n, int steps, float* pos4, float* vel4, float eps2, float damping,
                float dt, int nthreads)
{
    float* f = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    if (!f)
        return -1;
    for (int s = 0; s < steps; ++s) {
        orc_gravitation_rows_f32(n, pos4, eps2, 0, n, f, nthreads);
        orc_integrate_f32(n, pos4, vel4, f, dt, damping);
    }
    free(f);
    return 0;
}

/* FP64 rows [i0, i1) of orc_gravitation_f64 on worker threads (the same
 * partition rule as orc_gravitation_rows_f32); every row's sum is the serial
 * one, so the result does not depend on the thread count. */
typedef struct {
    int n, i0, i1;
    const double* pos4;
    double eps2;
    double* force3;
} rows64_job;

static void* rows64_worker(void* p)
{
    rows64_job* jb = (rows64_job*)p;
    grav_rows_f64(jb->n, jb->pos4, jb->eps2, jb->i0, jb->i1, jb->force3);
    return NULL;
}

static void grav_f64_threaded(int n, const double* pos4, double eps2, double* force3, int nthreads)
{
    if (nthreads <= 1 || n < nthreads) {
        grav_rows_f64(n, pos4, eps2, 0, n, force3);
        return;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    rows64_job* jobs = (rows64_job*)malloc(sizeof(rows64_job) * (size_t)nthreads);
    int per = n / nthreads, extra = n % nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int local = t < extra ? per + 1 : per;
        int start = t < extra ? t * (per + 1) : t * per + extra;
        jobs[t] = (rows64_job){n, start, start + local, pos4, eps2, force3};
        pthread_create(&th[t], NULL, rows64_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t)
        pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

int orc_run_f64_mt(int n, int steps, double* pos4, double* vel4, double eps2, double damping,
                   double dt, int nthreads)
{
    double* f = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    if (!f)
        return -1;
    for (int s = 0; s < steps; ++s) {
        grav_f64_threaded(n, pos4, eps2, f, nthreads);
        orc_integrate_f64(n, pos4, vel4, f, dt, damping);
    }
    free(f);
    return 0;
}

int orc_run_f64(int n, int steps, double* pos4, double* vel4, double eps2, double damping,
                double dt)
{
    return orc_run_f64_mt(n, steps, pos4, vel4, eps2, damping, dt, 1);
}

/* ------------------------------------------------------------------------ */
/* Initial conditions, seq/nbody.h:43-250 (vector helpers :43-85, generator
 * :89-250).  All arithmetic is float, as in the reference; values are stored
 * into T.  The random stream is glibc rand(), which the reference calls
 * without seeding (default seed 1).                                         */
void orc_srand(unsigned seed) { srand(seed); }

static inline float urand_f(void) { return rand() / (float)RAND_MAX; }

/* normalize(), seq/nbody.h:54-67: divide only when the length exceeds 1e-6
 * (compared in double, as the literal is double). */
static inline float normalize3(float v[3])
{
    float dist = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if ((double)dist > 1e-6) {
        v[0] /= dist;
        v[1] /= dist;
        v[2] /= dist;
    }
    return dist;
}

static inline float dot3(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

#define DEFINE_RANDOMIZE(NAME, T)                                                        \
    void NAME(int config, T* pos, T* vel, float* color, float clusterScale,              \
              float velocityScale, int numBodies, int vec4vel)                           \
    {                                                                                    \
        int p = 0, v = 0, i = 0;                                                         \
        if (config == 1) { /* SHELL, seq/nbody.h:141-194 */                              \
            float scale = clusterScale;                                                  \
            float vscale = scale * velocityScale;                                        \
            float inner = 2.5f * scale;                                                  \
            float outer = 4.0f * scale;                                                  \
            while (i < numBodies) {                                                      \
                float pt[3];                                                             \
                pt[0] = urand_f() * 2 - 1;                                               \
                pt[1] = urand_f() * 2 - 1;                                               \
                pt[2] = urand_f() * 2 - 1;                                               \
                float len = normalize3(pt);                                              \
                if (len > 1)                                                             \
                    continue;                                                            \
                pos[p++] = pt[0] * (inner + (outer - inner) * (float)rand() / (float)RAND_MAX); \
                pos[p++] = pt[1] * (inner + (outer - inner) * (float)rand() / (float)RAND_MAX); \
                pos[p++] = pt[2] * (inner + (outer - inner) * (float)rand() / (float)RAND_MAX); \
                pos[p++] = 1.0f;                                                         \
                float axis[3] = {0.0f, 0.0f, 1.0f};                                      \
                normalize3(axis);                                                        \
                if ((double)(1 - dot3(pt, axis)) < 1e-6) {                               \
                    axis[0] = pt[1];                                                     \
                    axis[1] = pt[0];                                                     \
                    normalize3(axis);                                                    \
                }                                                                        \
                float vv[3] = {(float)pos[4 * i], (float)pos[4 * i + 1], (float)pos[4 * i + 2]}; \
                float cr[3];                                                             \
                cr[0] = vv[1] * axis[2] - vv[2] * axis[1];                               \
                cr[1] = vv[2] * axis[0] - vv[0] * axis[2];                               \
                cr[2] = vv[0] * axis[1] - vv[1] * axis[0];                               \
                vel[v++] = cr[0] * vscale;                                               \
                vel[v++] = cr[1] * vscale;                                               \
                vel[v++] = cr[2] * vscale;                                               \
                if (vec4vel)                                                             \
                    vel[v++] = 1.0f;                                                     \
                i++;                                                                     \
            }                                                                            \
        } else if (config == 2) { /* EXPAND, seq/nbody.h:196-233 */                      \
            float scale = clusterScale * (float)numBodies / (1024.f);                    \
            if (scale < 1.0f)                                                            \
                scale = clusterScale;                                                    \
            float vscale = scale * velocityScale;                                        \
            while (i < numBodies) {                                                      \
                float pt[3];                                                             \
                pt[0] = urand_f() * 2 - 1;                                               \
                pt[1] = urand_f() * 2 - 1;                                               \
                pt[2] = urand_f() * 2 - 1;                                               \
                float lenSqr = dot3(pt, pt);                                             \
                if (lenSqr > 1)                                                          \
                    continue;                                                            \
                pos[p++] = pt[0] * scale;                                                \
                pos[p++] = pt[1] * scale;                                                \
                pos[p++] = pt[2] * scale;                                                \
                pos[p++] = 1.0f;                                                         \
                vel[v++] = pt[0] * vscale;                                               \
                vel[v++] = pt[1] * vscale;                                               \
                vel[v++] = pt[2] * vscale;                                               \
                if (vec4vel)                                                             \
                    vel[v++] = 1.0f;                                                     \
                i++;                                                                     \
            }                                                                            \
        } else { /* RANDOM (and default), seq/nbody.h:95-139 */                          \
            float q = (float)numBodies / (1024.0f);                                      \
            float scale = clusterScale * (q > 1.0f ? q : 1.0f);                          \
            float vscale = velocityScale * scale;                                        \
            while (i < numBodies) {                                                      \
                float pt[3], ve[3];                                                      \
                pt[0] = urand_f() * 2 - 1;                                               \
                pt[1] = urand_f() * 2 - 1;                                               \
                pt[2] = urand_f() * 2 - 1;                                               \
                if (dot3(pt, pt) > 1)                                                    \
                    continue;                                                            \
                ve[0] = urand_f() * 2 - 1;                                               \
                ve[1] = urand_f() * 2 - 1;                                               \
                ve[2] = urand_f() * 2 - 1;                                               \
                if (dot3(ve, ve) > 1)                                                    \
                    continue;                                                            \
                pos[p++] = pt[0] * scale;                                                \
                pos[p++] = pt[1] * scale;                                                \
                pos[p++] = pt[2] * scale;                                                \
                pos[p++] = 1.0f;                                                         \
                vel[v++] = ve[0] * vscale;                                               \
                vel[v++] = ve[1] * vscale;                                               \
                vel[v++] = ve[2] * vscale;                                               \
                if (vec4vel)                                                             \
                    vel[v++] = 1.0f;                                                     \
                i++;                                                                     \
            }                                                                            \
        }                                                                                \
        if (color) { /* seq/nbody.h:236-248 */                                           \
            int c = 0;                                                                   \
            for (int k = 0; k < numBodies; k++) {                                        \
                color[c++] = urand_f();                                                  \
                color[c++] = urand_f();                                                  \
                color[c++] = urand_f();                                                  \
                color[c++] = 1.0f;                                                       \
            }                                                                            \
        }                                                                                \
    }

DEFINE_RANDOMIZE(orc_randomize_f32, float)
DEFINE_RANDOMIZE(orc_randomize_f64, double)

/* demoParams and the N-dependent scales, seq/nbody_main.cc:57-66,112-144.
 * Fields: timestep, clusterScale, velocityScale, softening, damping. */
void orc_demo_params(int preset, int n, float out[5])
{
    static const float demo[7][5] = {
        {0.016f, 1.54f, 8.0f, 0.1f, 1.0f},      {0.016f, 0.68f, 20.0f, 0.1f, 1.0f},
        {0.0006f, 0.16f, 1000.0f, 1.0f, 1.0f},  {0.0006f, 0.16f, 1000.0f, 1.0f, 1.0f},
        {0.0019f, 0.32f, 276.0f, 1.0f, 1.0f},   {0.0016f, 0.32f, 272.0f, 0.145f, 1.0f},
        {0.016000f, 6.040000f, 0.000000f, 1.000000f, 1.000000f},
    };
    if (preset > 6 || preset < 0)
        preset = 0;
    memcpy(out, demo[preset], sizeof(float) * 5);
    if (n <= 1024) {
        out[1] = 1.52f; out[2] = 2.f;
    } else if (n <= 2048) {
        out[1] = 1.56f; out[2] = 2.64f;
    } else if (n <= 4096) {
        out[1] = 1.68f; out[2] = 2.98f;
    } else if (n <= 8192) {
        out[1] = 1.98f; out[2] = 2.9f;
    } else if (n <= 16384) {
        out[1] = 1.54f; out[2] = 8.f;
    } else if (n <= 32768) {
        out[1] = 1.44f; out[2] = 11.f;
    }
}

/* ------------------------------------------------------------------------ */
/* Frame file, seq/nbody_main.cc:195-211: per frame the label
 * frame_index*timestep ("%f"), then one "%.5f %.5f %.5f %.5f" line per body. */
#define DEFINE_WRITE_FRAMES(NAME, T)                                                     \
    int NAME(const char* path, const T* frames, int nframes, int n, float timestep)      \
    {                                                                                    \
        FILE* fp = fopen(path, "w");                                                     \
        if (!fp)                                                                         \
            return -1;                                                                   \
        for (int f = 0; f < nframes; f++) {                                              \
            fprintf(fp, "%f\n", f * timestep);                                           \
            const T* b = frames + (size_t)f * n * 4;                                     \
            for (int j = 0; j < n; j++) {                                                \
                fprintf(fp, "%.5f ", (double)b[j * 4 + 0]);                              \
                fprintf(fp, "%.5f ", (double)b[j * 4 + 1]);                              \
                fprintf(fp, "%.5f ", (double)b[j * 4 + 2]);                              \
                fprintf(fp, "%.5f\n", (double)b[j * 4 + 3]);                             \
            }                                                                            \
        }                                                                                \
        fclose(fp);                                                                      \
        return 0;                                                                        \
    }

DEFINE_WRITE_FRAMES(orc_write_frames_f32, float)
DEFINE_WRITE_FRAMES(orc_write_frames_f64, double)

/* ------------------------------------------------------------------------ */
/* FP64 physical mode.  No reference counterpart: the reference's closest
 * analog is BodySystem<double> (seq/bodysystem.cc:184-185), whose plain
 * summation is restated above.  Here every per-body sum is compensated
 * (Neumaier), so the result is accurate to a few FP64 ulps of the largest
 * term and serves as the error yardstick for the FP32 and mixed paths.
 * The self pair is skipped: for eps2 > 0 it contributes exactly zero (r = 0),
 * and skipping it keeps the functions finite at eps2 = 0.                   */
typedef struct {
    double s, c;
} ksum;

static inline void ksum_add(ksum* k, double x)
{
    double t = k->s + x;
    if (fabs(k->s) >= fabs(x))
        k->c += (k->s - t) + x;
    else
        k->c += (x - t) + k->s;
    k->s = t;
}

static inline double ksum_val(const ksum* k) { return k->s + k->c; }

/* The compensated FP64 field on body i (three values into out3). */
static void accel_row_f64(int n, const double* pos4, double eps2, int i, double* out3)
{
    const double* pi = pos4 + 4 * (size_t)i;
    ksum ax = {0, 0}, ay = {0, 0}, az = {0, 0};
    for (int j = 0; j < n; ++j) {
        if (j == i)
            continue;
        const double* pj = pos4 + 4 * (size_t)j;
        double rx = pj[0] - pi[0], ry = pj[1] - pi[1], rz = pj[2] - pi[2];
        double d2 = rx * rx + ry * ry + rz * rz + eps2;
        double inv = 1.0 / sqrt(d2);
        double s = pj[3] * (inv * inv * inv);
        ksum_add(&ax, rx * s);
        ksum_add(&ay, ry * s);
        ksum_add(&az, rz * s);
    }
    out3[0] = ksum_val(&ax);
    out3[1] = ksum_val(&ay);
    out3[2] = ksum_val(&az);
}

void orc_accel_f64(int n, const double* pos4, double eps2, double* acc3)
{
    for (int i = 0; i < n; ++i)
        accel_row_f64(n, pos4, eps2, i, acc3 + 3 * (size_t)i);
}

/* Rows rows[0..nrows) of orc_accel_f64, threaded with the reference partition
 * rule over the row list (the first nrows%T workers take one row more). */
typedef struct {
    int n;
    const double* pos4;
    double eps2;
    const int* rows;
    int k0, k1;
    double* acc3;
    int cpu;
} accel_rows_job;

static void* accel_rows_worker(void* p)
{
    accel_rows_job* jb = (accel_rows_job*)p;
    if (jb->cpu >= 0) {
        cpu_set_t set;
        CPU_ZERO(&set);
        CPU_SET(jb->cpu, &set);
        (void)pthread_setaffinity_np(pthread_self(), sizeof(set), &set); /* best effort */
    }
    for (int k = jb->k0; k < jb->k1; ++k)
        accel_row_f64(jb->n, jb->pos4, jb->eps2, jb->rows[k], jb->acc3 + 3 * (size_t)k);
    return NULL;
}

int orc_accel_rows_f64(int n, const double* pos4, double eps2, int nrows, const int* rows,
                       double* acc3, int nthreads)
{
    for (int k = 0; k < nrows; ++k)
        if (rows[k] < 0 || rows[k] >= n)
            return -1;
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > nrows)
        nthreads = nrows > 0 ? nrows : 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    accel_rows_job* jobs = (accel_rows_job*)malloc(sizeof(accel_rows_job) * (size_t)nthreads);
    int per = nrows / nthreads, extra = nrows % nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int local = t < extra ? per + 1 : per;
        int start = t < extra ? t * (per + 1) : t * per + extra;
        jobs[t] = (accel_rows_job){n, pos4, eps2, rows, start, start + local, acc3,
                                   nthreads > 1 ? allowed_cpu(t) : -1};
        if (nthreads > 1)
            pthread_create(&th[t], NULL, accel_rows_worker, &jobs[t]);
        else
            accel_rows_worker(&jobs[t]);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; ++t)
            pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    return 0;
}

void orc_accel_ld(int n, const double* pos4, double eps2, double* acc3)
{
    for (int i = 0; i < n; ++i) {
        const double* pi = pos4 + 4 * (size_t)i;
        long double ax = 0, ay = 0, az = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const double* pj = pos4 + 4 * (size_t)j;
            long double rx = (long double)pj[0] - pi[0];
            long double ry = (long double)pj[1] - pi[1];
            long double rz = (long double)pj[2] - pi[2];
            long double d2 = rx * rx + ry * ry + rz * rz + eps2;
            long double s = (long double)pj[3] / (d2 * sqrtl(d2));
            ax += rx * s;
            ay += ry * s;
            az += rz * s;
        }
        acc3[3 * (size_t)i + 0] = (double)ax;
        acc3[3 * (size_t)i + 1] = (double)ay;
        acc3[3 * (size_t)i + 2] = (double)az;
    }
}

void orc_step_f64(int n, double* pos4, double* vel4, double eps2, double dt, double damping,
                  int kick)
{
    double* a = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    orc_accel_f64(n, pos4, eps2, a);
    for (int i = 0; i < n; ++i) {
        double* p = pos4 + 4 * (size_t)i;
        double* v = vel4 + 4 * (size_t)i;
        double w = kick == 0 ? p[3] : 1.0;
        for (int k = 0; k < 3; ++k) {
            double u = (v[k] + (a[3 * (size_t)i + k] * w) * dt) * damping;
            v[k] = u;
            p[k] += u * dt;
        }
    }
    free(a);
}

void orc_energy_f64(int n, const double* pos4, const double* vel4, double eps2,
                    double* kinetic, double* potential, double* momentum3)
{
    ksum K = {0, 0}, W = {0, 0}, P[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int i = 0; i < n; ++i) {
        const double* p = pos4 + 4 * (size_t)i;
        const double* v = vel4 + 4 * (size_t)i;
        double m = p[3];
        ksum_add(&K, 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        for (int k = 0; k < 3; ++k)
            ksum_add(&P[k], m * v[k]);
        for (int j = i + 1; j < n; ++j) {
            const double* q = pos4 + 4 * (size_t)j;
            double rx = q[0] - p[0], ry = q[1] - p[1], rz = q[2] - p[2];
            ksum_add(&W, -m * q[3] / sqrt(rx * rx + ry * ry + rz * rz + eps2));
        }
    }
    *kinetic = ksum_val(&K);
    *potential = ksum_val(&W);
    for (int k = 0; k < 3; ++k)
        momentum3[k] = ksum_val(&P[k]);
}

static double loss_of(int n, const double* pos4, const double* target3)
{
    ksum L = {0, 0};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            double d = pos4[4 * (size_t)i + k] - target3[3 * (size_t)i + k];
            ksum_add(&L, d * d);
        }
    return ksum_val(&L) / n;
}

double orc_loss_f64(int n, int steps, const double* pos4, const double* vel4,
                    const double* target3, double eps2, double dt, double damping, int kick)
{
    double* x = (double*)malloc(sizeof(double) * 4 * (size_t)n);
    double* v = (double*)malloc(sizeof(double) * 4 * (size_t)n);
    memcpy(x, pos4, sizeof(double) * 4 * (size_t)n);
    memcpy(v, vel4, sizeof(double) * 4 * (size_t)n);
    for (int s = 0; s < steps; ++s)
        orc_step_f64(n, x, v, eps2, dt, damping, kick);
    double L = loss_of(n, x, target3);
    free(x);
    free(v);
    return L;
}

/* Reverse mode.  Forward step:  a_i = sum_j m_j K(r_ij), K(r) = r d^-3,
 * d^2 = |r|^2 + eps2, r_ij = x_j - x_i;  u_i = D (v_i + k_i a_i dt);
 * x_i' = x_i + u_i dt;  v_i' = u_i;  k_i = m_i (kick 0) or 1 (kick 1).
 * Adjoint of one step, given xb' and vb':
 *   ub = vb' + xb' dt;  vb = D ub;  alpha_i = D k_i dt ub_i;
 *   mb_i += D dt ub_i . a_i                       (kick 0 only)
 *   xb_k = xb'_k + sum_j J(r_kj) (m_k alpha_j - m_j alpha_k),
 *          J(r) = d^-3 I - 3 d^-5 r r^T
 *   mb_k += - sum_j alpha_j . K(r_kj).
 * All states are stored (the checkpointed GPU version recomputes them). */
double orc_adjoint_f64(int n, int steps, const double* pos4, const double* vel4,
                       const double* target3, double eps2, double dt, double damping,
                       int kick, double* dx0, double* dv0, double* dm)
{
    size_t n4 = 4 * (size_t)n, n3 = 3 * (size_t)n;
    double* xs = (double*)malloc(sizeof(double) * n4 * (size_t)(steps + 1));
    double* v = (double*)malloc(sizeof(double) * n4);
    double* a = (double*)malloc(sizeof(double) * n3);
    double* xb = (double*)malloc(sizeof(double) * n3);
    double* vb = (double*)malloc(sizeof(double) * n3);
    double* al = (double*)malloc(sizeof(double) * n3);
    double* nx = (double*)malloc(sizeof(double) * n3);
    memcpy(xs, pos4, sizeof(double) * n4);
    memcpy(v, vel4, sizeof(double) * n4);
    for (int s = 0; s < steps; ++s) {
        memcpy(xs + n4 * (size_t)(s + 1), xs + n4 * (size_t)s, sizeof(double) * n4);
        orc_step_f64(n, xs + n4 * (size_t)(s + 1), v, eps2, dt, damping, kick);
    }
    const double* xT = xs + n4 * (size_t)steps;
    double L = loss_of(n, xT, target3);
    for (int i = 0; i < n; ++i) {
        dm[i] = 0.0;
        for (int k = 0; k < 3; ++k) {
            xb[3 * i + k] = 2.0 * (xT[4 * i + k] - target3[3 * i + k]) / n;
            vb[3 * i + k] = 0.0;
        }
    }
    for (int s = steps - 1; s >= 0; --s) {
        const double* x = xs + n4 * (size_t)s;
        orc_accel_f64(n, x, eps2, a);
        for (int i = 0; i < n; ++i) {
            double ki = kick == 0 ? x[4 * i + 3] : 1.0;
            double dota = 0.0;
            for (int k = 0; k < 3; ++k) {
                double ub = vb[3 * i + k] + xb[3 * i + k] * dt;
                vb[3 * i + k] = damping * ub;
                al[3 * i + k] = damping * ki * dt * ub;
                dota += ub * a[3 * i + k];
            }
            if (kick == 0)
                dm[i] += damping * dt * dota;
        }
        for (int i = 0; i < n; ++i) {
            const double* pi = x + 4 * (size_t)i;
            double mi = pi[3];
            const double* ai = al + 3 * (size_t)i;
            ksum gx = {0, 0}, gy = {0, 0}, gz = {0, 0}, gm = {0, 0};
            for (int j = 0; j < n; ++j) {
                if (j == i)
                    continue;
                const double* pj = x + 4 * (size_t)j;
                const double* aj = al + 3 * (size_t)j;
                double rx = pj[0] - pi[0], ry = pj[1] - pi[1], rz = pj[2] - pi[2];
                double d2 = rx * rx + ry * ry + rz * rz + eps2;
                double inv = 1.0 / sqrt(d2);
                double inv3 = inv * inv * inv;
                double inv5 = inv3 * inv * inv;
                double wx = mi * aj[0] - pj[3] * ai[0];
                double wy = mi * aj[1] - pj[3] * ai[1];
                double wz = mi * aj[2] - pj[3] * ai[2];
                double rw = rx * wx + ry * wy + rz * wz;
                ksum_add(&gx, inv3 * wx - 3.0 * inv5 * rx * rw);
                ksum_add(&gy, inv3 * wy - 3.0 * inv5 * ry * rw);
                ksum_add(&gz, inv3 * wz - 3.0 * inv5 * rz * rw);
                ksum_add(&gm, -(aj[0] * rx + aj[1] * ry + aj[2] * rz) * inv3);
            }
            nx[3 * i + 0] = xb[3 * i + 0] + ksum_val(&gx);
            nx[3 * i + 1] = xb[3 * i + 1] + ksum_val(&gy);
            nx[3 * i + 2] = xb[3 * i + 2] + ksum_val(&gz);
            dm[i] += ksum_val(&gm);
        }
        memcpy(xb, nx, sizeof(double) * n3);
    }
    memcpy(dx0, xb, sizeof(double) * n3);
    memcpy(dv0, vb, sizeof(double) * n3);
    free(xs);
    free(v);
    free(a);
    free(xb);
    free(vb);
    free(al);
    free(nx);
    return L;
}
