// task.h -- the nbody driver's CPU tasks: initial conditions and frame output.
// Same names and runImpl signatures as benchapps/nbody/utc/task.h:86-97.
#pragma once

#include <cstdio>

#include "Utc.h"

using namespace iUtc;

enum NBodyConfig { NBODY_CONFIG_RANDOM, NBODY_CONFIG_SHELL, NBODY_CONFIG_EXPAND, NBODY_NUM_CONFIGS };

// Fills pos (x,y,z,m) and vel (vx,vy,vz[,w]) with the reference's initial
// conditions from the C library's rand() stream (benchapps/nbody/utc/task.cc:11-170).
template <typename T>
class NbodyInit : public UserTaskBase {
public:
    void runImpl(NBodyConfig config, T* pos, T* vel, float* color, float clusterScale,
                 float velocityScale, int numBodies, bool vec4vel);
};

// Appends one frame: "%f\n" label, then "%.5f %.5f %.5f %.5f\n" per body
// (benchapps/nbody/utc/task.cc:173-181).
template <typename T>
class Output : public UserTaskBase {
public:
    void runImpl(FILE** fp, T* buffer, float timestamp, int numBodies);
};
