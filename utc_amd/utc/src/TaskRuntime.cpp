// TaskRuntime.cpp -- worker threads and the ordered job queue (see TaskRuntime.h).
#include "TaskRuntime.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <exception>

#include "UserTaskBase.h"

namespace iUtc {

namespace {
thread_local TaskRuntime* tl_runtime = nullptr;
thread_local TaskBase* tl_task = nullptr;
}  // namespace

TaskRuntime* currentRuntime() { return tl_runtime; }
TaskBase* getCurrentTask() { return tl_task; }
void setCurrentTask(TaskBase* t, TaskRuntime* rt)
{
    tl_task = t;
    tl_runtime = rt;
}

TaskRuntime::TaskRuntime(std::string name, int nthreads, TaskType type, UserTaskBase* obj,
                         TaskBase* owner)
    : m_name(std::move(name)), m_nthreads(nthreads), m_type(type), m_obj(obj), m_owner(owner)
{
    m_obj->setTaskShape(nthreads, 1, 1, nthreads);
    for (int k = 0; k < nthreads; ++k)
        m_threads.emplace_back(&TaskRuntime::threadMain, this, k);
}

TaskRuntime::~TaskRuntime()
{
    if (!m_finished)
        finish();
}

void TaskRuntime::push(Job kind, std::function<void()> fn)
{
    m_jobs.push_back(Entry{kind, std::move(fn), m_nthreads});
}

void TaskRuntime::enqueue(Job kind, std::function<void()> fn)
{
    {
        std::lock_guard<std::mutex> lk(m_mu);
        push(kind, std::move(fn));
    }
    m_cv.notify_all();
}

size_t TaskRuntime::queuedJobs() const
{
    std::lock_guard<std::mutex> lk(m_mu);
    return m_jobs.size();
}

void TaskRuntime::wait()
{
    std::unique_lock<std::mutex> lk(m_mu);
    m_waitPending = m_nthreads;
    push(Job::wait, std::function<void()>());
    m_cv.notify_all();
    m_waitCv.wait(lk, [this] { return m_waitPending == 0; });
}

void TaskRuntime::finish()
{
    if (m_finished)
        return;
    enqueue(Job::finish, std::function<void()>());
    for (auto& t : m_threads)
        if (t.joinable())
            t.join();
    m_finished = true;
}

bool TaskRuntime::failed() const
{
    std::lock_guard<std::mutex> lk(m_mu);
    return !m_error.empty();
}

std::string TaskRuntime::error() const
{
    std::lock_guard<std::mutex> lk(m_mu);
    return m_error;
}

void TaskRuntime::barrier()
{
    std::unique_lock<std::mutex> lk(m_barMu);
    const unsigned long gen = m_barGen;
    if (++m_barCount == m_nthreads) {
        m_barCount = 0;
        ++m_barGen;
        m_barCv.notify_all();
        return;
    }
    m_barCv.wait(lk, [&] { return gen != m_barGen; });
}

void TaskRuntime::threadMain(int lrank)
{
    hipStream_t stream = nullptr;
    int device = -1;
    if (m_type == TaskType::gpu_task) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {
            device = lrank % ndev;  // src/gpu/CudaDeviceManager.cc:141-143
            if (hipSetDevice(device) != hipSuccess ||
                hipStreamCreateWithFlags(&stream, hipStreamDefault) != hipSuccess) {
                std::fprintf(stderr, "[%s] thread %d: cannot bind device %d\n", m_name.c_str(), lrank, device);
                stream = nullptr;
            }
        } else {
            std::fprintf(stderr, "[%s] thread %d: gpu task without a HIP device\n", m_name.c_str(), lrank);
        }
    }
    setCurrentTask(m_owner, this);
    m_obj->preInit(lrank, lrank, 0, stream, device);
    size_t idx = 0;
    for (;;) {
        std::unique_lock<std::mutex> lk(m_mu);
        m_cv.wait(lk, [&] { return m_base + m_jobs.size() > idx; });
        Entry& e = m_jobs[idx - m_base];
        const Job kind = e.kind;
        // the last worker to take a job moves its function out; the queue then
        // drops every leading job all workers have taken
        std::function<void()> fn = e.pending == 1 ? std::move(e.fn) : e.fn;
        if (--e.pending == 0) {
            while (!m_jobs.empty() && m_jobs.front().pending == 0) {
                m_jobs.pop_front();
                ++m_base;
            }
        }
        lk.unlock();
        ++idx;
        if (kind == Job::finish)
            break;
        if (kind == Job::wait) {
            std::lock_guard<std::mutex> g(m_mu);
            if (--m_waitPending == 0)
                m_waitCv.notify_all();
            continue;
        }
        try {
            if (fn)
                fn();
        } catch (const std::exception& e) {
            std::fprintf(stderr, "[%s] thread %d: exception in user task: %s\n", m_name.c_str(), lrank, e.what());
            std::lock_guard<std::mutex> g(m_mu);
            m_error = e.what();
        } catch (...) {
            std::fprintf(stderr, "[%s] thread %d: unknown exception in user task\n", m_name.c_str(), lrank);
            std::lock_guard<std::mutex> g(m_mu);
            m_error = "unknown exception";
        }
        if (kind == Job::init)
            barrier();  // src/gpu/TaskGPU.cc:293-296: init ends with a thread sync
    }
    m_obj->preExit();
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
    setCurrentTask(nullptr, nullptr);
}

}  // namespace iUtc
