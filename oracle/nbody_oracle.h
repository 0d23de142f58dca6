/*
 * nbody_oracle.h -- CPU oracle for the UTC nbody hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only the
 * tests, the smoke entry point and the CPU-baseline leg of the benchmark may
 * load it.  The HIP library never links or calls anything declared here.
 *
 * Reference semantics (paths relative to the reference's utc-0.1/ tree):
 *   pair interaction      benchapps/nbody/seq/bodysystem.h:70-96
 *   all-pairs gravitation benchapps/nbody/seq/bodysystem.cc:103-133
 *   damped kick-drift     benchapps/nbody/seq/bodysystem.cc:135-181
 *   step = force then integrate  benchapps/nbody/seq/bodysystem.cc:64-73
 *   initial conditions    benchapps/nbody/seq/nbody.h:43-250
 *   presets / N scaling   benchapps/nbody/seq/nbody_main.cc:57-66,109-144
 *   frame writer          benchapps/nbody/seq/nbody_main.cc:195-211
 *   body partition rule   benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:41-49
 *
 * Layouts follow the reference: pos = T[4N] (x,y,z,m), vel = T[4N]
 * (vx,vy,vz,w), force = T[3N].
 */
#ifndef UTCNB_ORACLE_H_
#define UTCNB_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference mode, bit-for-bit restatement of seq/ ------------------- */
void orc_gravitation_f32(int n, const float* pos4, float eps2, float* force3);
void orc_gravitation_f64(int n, const double* pos4, double eps2, double* force3);
/* rows [i0,i1) only, optionally threaded with the reference partition rule */
void orc_gravitation_rows_f32(int n, const float* pos4, float eps2, int i0, int i1,
                              float* force3, int nthreads);
void orc_integrate_f32(int n, float* pos4, float* vel4, const float* force3,
                       float dt, float damping);
void orc_integrate_f64(int n, double* pos4, double* vel4, const double* force3,
                       double dt, double damping);
/* `steps` calls of BodySystem<T>::update(dt); scratch force buffer is internal */
int orc_run_f32(int n, int steps, float* pos4, float* vel4, float eps2, float damping,
                float dt, int nthreads);
int orc_run_f64(int n, int steps, double* pos4, double* vel4, double eps2,
                double damping, double dt);
/* the same on nthreads worker threads (rows split by the reference partition
 * rule; bit-identical to orc_run_f64 for any thread count) */
int orc_run_f64_mt(int n, int steps, double* pos4, double* vel4, double eps2,
                   double damping, double dt, int nthreads);

/* ---- initial conditions (glibc rand(), as the reference) --------------- */
void orc_srand(unsigned seed);
/* config: 0 RANDOM, 1 SHELL, 2 EXPAND (seq/nbody.h:14-20) */
void orc_randomize_f32(int config, float* pos4, float* vel4, float* color4,
                       float clusterScale, float velocityScale, int n, int vec4vel);
void orc_randomize_f64(int config, double* pos4, double* vel4, float* color4,
                       float clusterScale, float velocityScale, int n, int vec4vel);
/* out = {timestep, clusterScale, velocityScale, softening, damping} */
void orc_demo_params(int preset, int n, float out[5]);

/* ---- frame file (text) -------------------------------------------------- */
int orc_write_frames_f32(const char* path, const float* frames, int nframes, int n,
                         float timestep);
int orc_write_frames_f64(const char* path, const double* frames, int nframes, int n,
                         float timestep);

/* ---- FP64 physical mode (no reference counterpart) ---------------------- */
/* a_i = sum_j m_j r_ij / (|r_ij|^2+eps2)^{3/2}, Neumaier-compensated sum over j */
void orc_accel_f64(int n, const double* pos4, double eps2, double* acc3);
/* same, summed in long double with the pair in long double (tiny-N brute force) */
void orc_accel_ld(int n, const double* pos4, double eps2, double* acc3);
/* rows rows[0..nrows) of orc_accel_f64 (bit for bit) into acc3[3*nrows],
 * threaded with the reference partition rule; -1 if a row is out of range */
int orc_accel_rows_f64(int n, const double* pos4, double eps2, int nrows, const int* rows,
                       double* acc3, int nthreads);
/* kick=0: v=(v+a*m_i*dt)*damping (reference kick factor), kick=1: v=(v+a*dt)*damping;
 * then x += v*dt.  State in FP64. */
void orc_step_f64(int n, double* pos4, double* vel4, double eps2, double dt,
                  double damping, int kick);
/* kinetic, potential (pairs i<j), total energy and momentum */
void orc_energy_f64(int n, const double* pos4, const double* vel4, double eps2,
                    double* kinetic, double* potential, double* momentum3);
/* L = (1/n) sum_i |x_i(T) - t_i|^2 after `steps` FP64 steps */
double orc_loss_f64(int n, int steps, const double* pos4, const double* vel4,
                    const double* target3, double eps2, double dt, double damping,
                    int kick);
/* reverse mode through `steps` FP64 steps for the loss above.
 * Outputs dL/dx0 (3n), dL/dv0 (3n), dL/dm (n). Returns L. */
double orc_adjoint_f64(int n, int steps, const double* pos4, const double* vel4,
                       const double* target3, double eps2, double dt, double damping,
                       int kick, double* dx0, double* dv0, double* dm);

#ifdef __cplusplus
}
#endif
#endif
