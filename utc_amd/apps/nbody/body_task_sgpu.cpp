// body_task_sgpu.cpp -- BodySystemSGPU<T> over the gfx950 library.
//
// runImpl follows benchapps/nbody/utc/sgpu/body_task_sgpu.cu:41-137: double
// buffered positions and a velocity buffer in GpuData of the requested
// MemType, one upload, then `loops` steps each followed by a stream
// synchronisation, a download every outInterval steps into
// outbuffer[(i/outInterval - 1) * 4N], and runtime[0..3] = total, kernel,
// copy-in, copy-out seconds.  The kernel launch is the C ABI entry
// utcnb_integrate_f32/_f64 on the worker thread's stream.  Errors are printed
// and the run continues, as the reference does (benchapps/common/helper_err.h:11-18).
#include "body_task_sgpu.h"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

inline void report(int rc, const char* what)
{
    if (rc != UTCNB_OK)
        std::fprintf(stderr, "utcnb error %d in %s: %s\n", rc, what, utcnb_last_error());
}

inline int integrate(const float* po, float* pn, float* v, int n, float dt, float e2, float d, int bs,
                     void* st)
{
    return utcnb_integrate_f32(po, pn, v, n, dt, e2, d, bs, st);
}

inline int integrate(const double* po, double* pn, double* v, int n, double dt, double e2, double d,
                     int bs, void* st)
{
    return utcnb_integrate_f64(po, pn, v, n, dt, e2, d, bs, st);
}

template <typename T>
constexpr int dtype_of() { return sizeof(T) == 4 ? UTCNB_F32 : UTCNB_F64; }

}  // namespace

template <typename T>
void BodySystemSGPU<T>::initImpl(unsigned int numBodies, T softeningSquared, T damping, T* pos,
                                 T* vel)
{
    if (__localThreadId == 0) {
        m_numBodies = numBodies;
        m_softeningSquared = softeningSquared;
        m_damping = damping;
        m_pos = pos;  // borrowed until wait()
        m_vel = vel;
    }
    intra_Barrier();
}

template <typename T>
void BodySystemSGPU<T>::setPrecision(int precision)
{
    if (__localThreadId == 0)
        m_precision = precision;
    intra_Barrier();
}

template <typename T>
void BodySystemSGPU<T>::runImpl(double* runtime, int loops, int outInterval, int blocksize,
                                T deltaTime, T* outbuffer, MemType memtype)
{
    if (m_precision != UTCNB_PREC_REFERENCE) {
        runFast(runtime, loops, outInterval, deltaTime, outbuffer, memtype);
        return;
    }
    const int n = (int)m_numBodies;
    Timer timer, timer0;
    timer0.start();
    GpuData<T> posBuffer[2] = {GpuData<T>((unsigned long)n * 4, memtype),
                               GpuData<T>((unsigned long)n * 4, memtype)};
    GpuData<T> velBuffer((unsigned long)n * 4, memtype);
    posBuffer[0].initH(m_pos);
    velBuffer.initH(m_vel);

    timer.start();
    posBuffer[0].sync();
    velBuffer.sync();
    double copyinTime = timer.stop();

    double kernelTime = 0, copyoutTime = 0;
    int cur = 0;
    for (int i = 0; i < loops;) {
        timer.start();
        report(integrate(posBuffer[cur].getD(), posBuffer[1 - cur].getD(true), velBuffer.getD(true), n,
                         deltaTime, m_softeningSquared, m_damping, blocksize, (void*)__streamId),
               "utcnb_integrate");
        utcHipCheck(hipStreamSynchronize(__streamId), "hipStreamSynchronize");
        kernelTime += timer.stop();
        ++i;
        if (i % outInterval == 0) {
            timer.start();
            posBuffer[1 - cur].sync();
            copyoutTime += timer.stop();
            posBuffer[1 - cur].fetch(outbuffer + (size_t)(i / outInterval - 1) * n * 4);
        }
        cur = 1 - cur;
    }
    runtime[0] = timer0.stop();
    runtime[1] = kernelTime;
    runtime[2] = copyinTime;
    runtime[3] = copyoutTime;
}

// The fast precisions keep their state inside a library handle (padded,
// possibly FP64); host transfers go through GpuData host buffers of the
// requested MemType.
template <typename T>
void BodySystemSGPU<T>::runFast(double* runtime, int loops, int outInterval, T deltaTime,
                                T* outbuffer, MemType memtype)
{
    const int n = (int)m_numBodies;
    Timer timer, timer0;
    timer0.start();
    utcnb_config c{};
    c.n = n;
    c.precision = m_precision;
    c.kick = UTCNB_KICK_REFERENCE;
    c.flags = UTCNB_FLAG_NO_NEWTON;
    c.eps2 = (double)m_softeningSquared;
    c.damping = (double)m_damping;
    c.dt = (double)deltaTime;
    utcnb_system* h = nullptr;
    int rc = utcnb_create(&c, (void*)__streamId, &h);
    report(rc, "utcnb_create");
    if (rc)
        return;
    GpuData<T> host((unsigned long)n * 4, memtype);
    GpuData<T> hvel((unsigned long)n * 4, memtype);
    host.initH(m_pos);
    hvel.initH(m_vel);
    timer.start();
    report(utcnb_set_state(h, host.getH(), hvel.getH(), dtype_of<T>(), UTCNB_HOST), "utcnb_set_state");
    report(utcnb_sync(h), "utcnb_sync");
    double copyinTime = timer.stop();
    double kernelTime = 0, copyoutTime = 0;
    for (int i = 0; i < loops;) {
        timer.start();
        report(utcnb_step(h, 1), "utcnb_step");
        report(utcnb_sync(h), "utcnb_sync");
        kernelTime += timer.stop();
        ++i;
        if (i % outInterval == 0) {
            timer.start();
            report(utcnb_get_state(h, host.getH(), nullptr, dtype_of<T>(), UTCNB_HOST), "utcnb_get_state");
            copyoutTime += timer.stop();
            std::memcpy(outbuffer + (size_t)(i / outInterval - 1) * n * 4, host.getH(),
                        sizeof(T) * 4 * (size_t)n);
        }
    }
    utcnb_destroy(h);
    runtime[0] = timer0.stop();
    runtime[1] = kernelTime;
    runtime[2] = copyinTime;
    runtime[3] = copyoutTime;
}

template class BodySystemSGPU<float>;
template class BodySystemSGPU<double>;
