// Timer.cpp -- see Timer.h.
#include "Timer.h"

namespace iUtc {

Timer::Timer(Timer_unit tu)
{
    switch (tu) {
    case SECONDS: m_ratio = 1.0; break;
    case MILLISEC: m_ratio = 1e3; break;
    case MICROSEC: m_ratio = 1e6; break;
    case NANOSEC: m_ratio = 1e9; break;
    }
    start();
}

void Timer::start()
{
    m_begin = std::chrono::steady_clock::now();
    clock_gettime(CLOCK_THREAD_CPUTIME_ID, &m_cpu_start);
}

void Timer::start(TimerValue& tv_out)
{
    start();
    tv_out = m_begin;
}

double Timer::stop()
{
    TimerValue now;
    return stop(now);
}

double Timer::stop(TimerValue& tv_stop)
{
    tv_stop = std::chrono::steady_clock::now();
    m_realTimeperiod = std::chrono::duration<double>(tv_stop - m_begin).count() * m_ratio;
    struct timespec c;
    clock_gettime(CLOCK_THREAD_CPUTIME_ID, &c);
    m_cpuTimeperiod = ((double)(c.tv_sec - m_cpu_start.tv_sec) +
                       (double)(c.tv_nsec - m_cpu_start.tv_nsec) * 1e-9) * m_ratio;
    return m_realTimeperiod;
}

double Timer::stop(TimerValue& tv_start, TimerValue& tv_stop)
{
    tv_stop = std::chrono::steady_clock::now();
    m_realTimeperiod = std::chrono::duration<double>(tv_stop - tv_start).count() * m_ratio;
    return m_realTimeperiod;
}

}  // namespace iUtc
