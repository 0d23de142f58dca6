// body_task_mgpu.cpp -- BodySystemMGPU<T> over the gfx950 library.
//
// runImpl follows benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:55-177:
// each thread keeps a GpuData copy of all N positions and of its own slice's
// new positions and velocities; per step it uploads the current positions
// from the shared host buffer, advances its slice, downloads the slice into
// the other host buffer and meets the other threads at a barrier; every
// outInterval steps it copies its slice of the frame into outbuffer.
// runtime[t][0..3] = total, kernel, copy-in, copy-out seconds of thread t.
// Errors are printed and the run continues (benchapps/common/helper_err.h:11-18).
#include "body_task_mgpu.h"

#include <cstdio>
#include <cstring>

namespace {

inline void report(int rc, const char* what)
{
    if (rc != UTCNB_OK)
        std::fprintf(stderr, "utcnb error %d in %s: %s\n", rc, what, utcnb_last_error());
}

inline int integrate_slice(const float* po, float* pn, float* v, int n, int start, int count, float dt,
                           float e2, float d, int bs, void* st)
{
    return utcnb_integrate_slice_f32(po, pn, v, n, start, count, dt, e2, d, bs, st);
}

inline int integrate_slice(const double* po, double* pn, double* v, int n, int start, int count, double dt,
                           double e2, double d, int bs, void* st)
{
    return utcnb_integrate_slice_f64(po, pn, v, n, start, count, dt, e2, d, bs, st);
}

}  // namespace

template <typename T>
thread_local int BodySystemMGPU<T>::local_numBodies;
template <typename T>
thread_local int BodySystemMGPU<T>::local_startBodyIndex;

template <typename T>
void BodySystemMGPU<T>::initImpl(unsigned int numBodies, T softeningSquared, T damping, T* pos, T* vel)
{
    if (__localThreadId == 0) {
        m_numBodies = numBodies;
        m_softeningSquared = softeningSquared;
        m_damping = damping;
        m_pos = pos;  // borrowed until wait()
        m_vel = vel;
        oldPosBuffer = new T[(size_t)numBodies * 4];
        newPosBuffer = new T[(size_t)numBodies * 4];
        std::memcpy(oldPosBuffer, m_pos, sizeof(T) * 4 * (size_t)numBodies);
    }
    intra_Barrier();
    // the reference's partition: the first N % T threads take one body more
    const int per = (int)numBodies / __numLocalThreads, extra = (int)numBodies % __numLocalThreads;
    if (__localThreadId < extra) {
        local_numBodies = per + 1;
        local_startBodyIndex = __localThreadId * (per + 1);
    } else {
        local_numBodies = per;
        local_startBodyIndex = __localThreadId * per + extra;
    }
    intra_Barrier();
}

template <typename T>
void BodySystemMGPU<T>::runImpl(double runtime[][4], int loops, int outInterval, int blocksize, T deltaTime,
                                T* outbuffer, MemType memtype)
{
    const int n = (int)m_numBodies, start = local_startBodyIndex, count = local_numBodies;
    Timer timer, timer0;
    GpuData<T> posBuffer[2] = {GpuData<T>((unsigned long)n * 4, memtype),
                               GpuData<T>((unsigned long)(count > 0 ? count : 1) * 4, memtype)};
    GpuData<T> velBuffer((unsigned long)(count > 0 ? count : 1) * 4, memtype);
    if (count > 0)
        velBuffer.initH(m_vel + (size_t)start * 4);

    timer0.start();
    timer.start();
    velBuffer.sync();
    double copyinTime = timer.stop();
    double kernelTime = 0, copyoutTime = 0;
    T* tmp[2] = {oldPosBuffer, newPosBuffer};
    int cur = 0;
    for (int i = 0; i < loops;) {
        // all bodies' current positions to this thread's device
        timer.start();
        posBuffer[0].putD(tmp[cur]);
        copyinTime += timer.stop();

        timer.start();
        if (count > 0) {
            report(integrate_slice(posBuffer[0].getD(), posBuffer[1].getD(true), velBuffer.getD(true), n, start,
                                   count, deltaTime, m_softeningSquared, m_damping, blocksize, (void*)__streamId),
                   "utcnb_integrate_slice");
            utcHipCheck(hipStreamSynchronize(__streamId), "hipStreamSynchronize");
        }
        kernelTime += timer.stop();

        // this thread's slice of the new positions to the shared host buffer
        timer.start();
        if (count > 0)
            posBuffer[1].fetchD(tmp[1 - cur] + (size_t)start * 4);
        copyoutTime += timer.stop();

        intra_Barrier();
        ++i;
        if (i % outInterval == 0 && count > 0) {
            const size_t offset = (size_t)(i / outInterval - 1) * n * 4;
            std::memcpy(outbuffer + offset + (size_t)start * 4, tmp[1 - cur] + (size_t)start * 4,
                        sizeof(T) * 4 * (size_t)count);
        }
        cur = 1 - cur;
    }
    intra_Barrier();
    runtime[__localThreadId][0] = timer0.stop();
    runtime[__localThreadId][1] = kernelTime;
    runtime[__localThreadId][2] = copyinTime;
    runtime[__localThreadId][3] = copyoutTime;
    intra_Barrier();  // every thread is done with the shared buffers
    if (__localThreadId == 0) {
        delete[] oldPosBuffer;
        delete[] newPosBuffer;
        oldPosBuffer = newPosBuffer = nullptr;
    }
}

template class BodySystemMGPU<float>;
template class BodySystemMGPU<double>;
