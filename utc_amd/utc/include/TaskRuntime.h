// TaskRuntime.h -- the worker threads and job queue behind Task<T>.
//
// Semantics follow the reference (include/Task.inc:264-570,
// src/gpu/TaskGPU.cc:195-400): one std::thread per local entry of the
// ProcList; init/run/exec append a job and return at once; every worker runs
// every job in order; an init job ends with a barrier of the local threads;
// wait() appends a wait job and blocks until every worker has reached it;
// finish() (also run by the destructor) appends a finish job and joins.  A GPU
// task binds worker k to device k % device_count and gives it its own HIP
// stream (default priority).  An exception escaping a job is reported on
// stderr and recorded (failed()); it never blocks wait().  A job leaves the
// queue (with its bound arguments) once every worker has taken it, so a task
// that runs many jobs keeps only the ones some worker has not reached.
#pragma once

#include <hip/hip_runtime_api.h>

#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "TaskBase.h"

class UserTaskBase;

namespace iUtc {

class TaskRuntime {
public:
    enum class Job { init, run, exec, wait, finish };

    TaskRuntime(std::string name, int nthreads, TaskType type, UserTaskBase* obj, TaskBase* owner);
    ~TaskRuntime();

    void enqueue(Job kind, std::function<void()> fn);
    void wait();
    void finish();
    bool finished() const { return m_finished; }
    bool failed() const;
    std::string error() const;
    int numThreads() const { return m_nthreads; }
    // jobs still queued (not yet taken by every worker)
    size_t queuedJobs() const;

    // barrier among this task's workers (used by intra_Barrier)
    void barrier();

private:
    void threadMain(int lrank);

    std::string m_name;
    int m_nthreads;
    TaskType m_type;
    UserTaskBase* m_obj;
    TaskBase* m_owner;
    std::vector<std::thread> m_threads;

    mutable std::mutex m_mu;
    std::condition_variable m_cv;
    struct Entry {
        Job kind;
        std::function<void()> fn;
        int pending;  // workers that have not taken the job yet
    };
    std::deque<Entry> m_jobs;  // jobs [m_base, m_base + m_jobs.size())
    size_t m_base = 0;
    void push(Job kind, std::function<void()> fn);  // caller holds m_mu
    // wait latch
    int m_waitPending = 0;
    std::condition_variable m_waitCv;
    // reusable barrier
    std::mutex m_barMu;
    std::condition_variable m_barCv;
    int m_barCount = 0;
    unsigned long m_barGen = 0;
    bool m_finished = false;
    std::string m_error;
};

// The runtime of the calling worker thread (nullptr outside a task).
TaskRuntime* currentRuntime();
void setCurrentTask(TaskBase* t, TaskRuntime* rt);

}  // namespace iUtc
