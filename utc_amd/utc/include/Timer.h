// Timer.h -- wall-clock interval timer (steady clock).
// Surface of include/Timer.h:18-71.
#pragma once

#include <chrono>
#include <ctime>

namespace iUtc {

enum Timer_unit { SECONDS, MILLISEC, MICROSEC, NANOSEC };

class Timer {
public:
    typedef std::chrono::steady_clock::time_point TimerValue;
    explicit Timer(Timer_unit tu = SECONDS);
    void start();
    void start(TimerValue& tv_out);
    // elapsed time since the last start(), in the timer's unit
    double stop();
    double stop(TimerValue& tv_stop);
    double stop(TimerValue& tv_start, TimerValue& tv_stop);
    double getRealTime() const { return m_realTimeperiod; }
    double getThreadCpuTime() const { return m_cpuTimeperiod; }

    Timer(const Timer&) = delete;
    Timer& operator=(const Timer&) = delete;

private:
    TimerValue m_begin;
    double m_realTimeperiod = 0.0;
    double m_cpuTimeperiod = 0.0;
    double m_ratio = 1.0;
    struct timespec m_cpu_start {};
};

}  // namespace iUtc
