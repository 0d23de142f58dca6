// body_task_sgpu.h -- the nbody GPU user task.
//
// Same class name, template parameter and initImpl/runImpl signatures as the
// reference's BodySystemSGPU<T> (benchapps/nbody/utc/sgpu/body_task_sgpu.h:16-41),
// so a driver written against it (benchapps/nbody/utc/nbody_main.cc:209-223)
// runs unchanged.  The arithmetic comes from the gfx950 library through the C
// ABI of utcnb.h.  One extra member, setPrecision(), selects the library
// precision through Task<T>::exec; the default is the reference's arithmetic.
#pragma once

#include "Utc.h"
#include "UtcGpu.h"
#include "utcnb.h"

using namespace iUtc;

template <typename T>
class BodySystemSGPU : public UserTaskBase {
private:
    unsigned int m_numBodies = 0;
    T* m_pos = nullptr;
    T* m_vel = nullptr;
    T m_softeningSquared = 0;
    T m_damping = 1;
    int m_precision = UTCNB_PREC_REFERENCE;

    void runFast(double* runtime, int loops, int outInterval, T deltaTime, T* outbuffer,
                 MemType memtype);

public:
    void initImpl(unsigned int numBodies, T softeningSquared, T damping, T* pos, T* vel);

    void runImpl(double* runtime, int loops, int outInterval, int blocksize, T deltaTime,
                 T* outbuffer, MemType memtype);

    // UTCNB_PREC_REFERENCE (default), _FP32, _MIXED or _FP64
    void setPrecision(int precision);
};
