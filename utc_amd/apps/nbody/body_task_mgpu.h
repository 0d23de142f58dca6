// body_task_mgpu.h -- the nbody GPU user task with one worker thread per GPU.
//
// Same class name, members and initImpl/runImpl signatures as the reference's
// BodySystemMGPU<T> (benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.h:16-46),
// so its driver (benchapps/nbody/utc-snode-mgpu/nbody_main.cc:198-230) runs
// unchanged: thread t of the task's ProcList owns a slice of the bodies (the
// first N % T threads one body more) and advances it against all N bodies on
// its own device and stream; the slices meet in host buffers each step.  The
// per-slice kernel is the C ABI entry utcnb_integrate_slice_f32/_f64 (the
// reference's arithmetic, bit for bit).  With more threads than GPUs, threads
// share devices (thread t on device t % device_count).
#pragma once

#include <vector>

#include "Utc.h"
#include "UtcGpu.h"
#include "utcnb.h"

using namespace iUtc;

template <typename T>
class BodySystemMGPU : public UserTaskBase {
private:
    unsigned int m_numBodies = 0;
    T* m_pos = nullptr;
    T* m_vel = nullptr;

    T m_softeningSquared = 0;
    T m_damping = 1;
    T* oldPosBuffer = nullptr;
    T* newPosBuffer = nullptr;

    static thread_local int local_numBodies;
    static thread_local int local_startBodyIndex;

public:
    void initImpl(unsigned int numBodies, T softeningSquared, T damping, T* pos, T* vel);

    // runtime[t][0..3] = thread t's total, kernel, copy-in, copy-out seconds
    void runImpl(double runtime[][4], int loops, int outInterval, int blocksize, T deltaTime,
                 T* outbuffer, MemType memtype);
};
