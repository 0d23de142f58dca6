// ref_harness.cc -- C entry points over the reference's own sequential nbody.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's unmodified
// benchapps/nbody/seq/bodysystem.cc (and its headers bodysystem.h, nbody.h)
// where they lie in the reference tree; the output goes to oracle/_ref/ and
// is never committed.  It lets the tests check the C restatement in
// nbody_oracle.c against the reference itself, and lets the benchmark time the
// reference's CPU path as a baseline.  Nothing here is shipped.
#include <cstring>

#include "bodysystem.h"  // benchapps/nbody/seq/bodysystem.h
#include "nbody.h"       // benchapps/nbody/seq/nbody.h

template <typename T>
static void run_seq(int n, int steps, T* pos4, T* vel4, T softening, T damping, T dt)
{
    // Same call sequence as benchapps/nbody/seq/nbody_main.cc:163-190.
    BodySystem<T> sys(n);
    sys.setDamping(damping);
    sys.setSoftening(softening);
    sys.setPosArray(pos4);
    sys.setVelArray(vel4);
    for (int s = 0; s < steps; ++s)
        sys.update(dt);
    std::memcpy(pos4, sys.getPosArray(), sizeof(T) * 4 * (size_t)n);
    std::memcpy(vel4, sys.getVelArray(), sizeof(T) * 4 * (size_t)n);
}

extern "C" {

void ref_run_f32(int n, int steps, float* pos4, float* vel4, float softening, float damping,
                 float dt)
{
    run_seq<float>(n, steps, pos4, vel4, softening, damping, dt);
}

void ref_run_f64(int n, int steps, double* pos4, double* vel4, double softening,
                 double damping, double dt)
{
    run_seq<double>(n, steps, pos4, vel4, softening, damping, dt);
}

void ref_randomize_f32(int config, float* pos4, float* vel4, float* color4, float clusterScale,
                       float velocityScale, int n, int vec4vel)
{
    randomizeBodies<float>((NBodyConfig)config, pos4, vel4, color4, clusterScale, velocityScale,
                           n, vec4vel != 0);
}

void ref_randomize_f64(int config, double* pos4, double* vel4, float* color4,
                       float clusterScale, float velocityScale, int n, int vec4vel)
{
    randomizeBodies<double>((NBodyConfig)config, pos4, vel4, color4, clusterScale,
                            velocityScale, n, vec4vel != 0);
}

void ref_srand(unsigned seed) { srand(seed); }

}  // extern "C"
