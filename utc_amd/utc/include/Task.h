// Task.h -- Task<T>: a user task of type T run by worker threads.
//
// Surface of the reference's include/Task.h:45-190 (construction from a
// ProcList and a TaskType; init, run, exec, wait, finish).  The user object is
// default constructed once and shared by all local worker threads
// (include/Task.inc:716-730); arguments of init/run/exec are copied into the
// job (std::bind), pointers are borrowed.
#pragma once

#include <functional>
#include <memory>
#include <string>

#include "ProcList.h"
#include "TaskBase.h"
#include "TaskRuntime.h"
#include "UserTaskBase.h"

namespace iUtc {

template <class T>
class Task : public TaskBase {
public:
    typedef T FunctionObjectType;

    Task() : Task(std::string("task"), ProcList(0), TaskType::cpu_task) {}
    explicit Task(ProcList rList) : Task(std::string("task"), rList, TaskType::cpu_task) {}
    Task(ProcList rList, TaskType tType) : Task(std::string("task"), rList, tType) {}
    explicit Task(std::string name) : Task(name, ProcList(0), TaskType::cpu_task) {}
    Task(std::string name, ProcList rList) : Task(name, rList, TaskType::cpu_task) {}
    Task(T* userTaskObj, ProcList rList, TaskType tType) { create("task", rList, tType, userTaskObj); }
    Task(std::string name, ProcList rList, TaskType tType) { create(name, rList, tType, nullptr); }

    ~Task()
    {
        if (m_rt && !m_rt->finished())
            m_rt->finish();
        m_rt.reset();
        if (!m_userSupplied)
            delete m_obj;
    }

    template <typename... Args>
    void init(Args... args)
    {
        if (!m_rt)
            return;
        m_rt->enqueue(TaskRuntime::Job::init, std::bind(&T::initImpl, m_obj, args...));
    }

    template <typename... Args>
    void run(Args... args)
    {
        if (!m_rt)
            return;
        m_rt->enqueue(TaskRuntime::Job::run, std::bind(&T::runImpl, m_obj, args...));
    }

    template <typename... Args>
    void exec(void (T::*user_fun)(Args...), Args... args)
    {
        if (!m_rt)
            return;
        m_rt->enqueue(TaskRuntime::Job::exec, std::bind(user_fun, m_obj, args...));
    }

    void wait()
    {
        if (m_rt)
            m_rt->wait();
    }

    void finish()
    {
        if (m_rt && !m_rt->finished())
            m_rt->finish();
    }

    bool failed() const { return m_rt && m_rt->failed(); }
    // jobs some local worker has not taken yet (not in the reference)
    size_t queuedJobs() const { return m_rt ? m_rt->queuedJobs() : 0; }
    std::string error() const { return m_rt ? m_rt->error() : std::string(); }
    T* getUserTaskObj() { return m_obj; }
    bool hasActiveLocalThread() const { return m_rt && !m_rt->finished(); }
    void waitLocalThreadFinish() { finish(); }

    Task(const Task&) = delete;
    Task& operator=(const Task&) = delete;

private:
    void create(const std::string& name, const ProcList& rList, TaskType tType, T* obj)
    {
        m_Name = name;
        m_TaskType = tType;
        int local = 0;
        for (int i = 0; i < rList.getNumProcs(); ++i)
            if (rList.getProc(i) == 0)  // this process
                ++local;
        m_numLocalThreads = local;
        m_userSupplied = obj != nullptr;
        if (local == 0)
            return;  // no thread of this task on this process: calls are no-ops
        m_obj = obj ? obj : new T();
        m_rt.reset(new TaskRuntime(name, local, tType, m_obj, this));
    }

    T* m_obj = nullptr;
    bool m_userSupplied = false;
    std::unique_ptr<TaskRuntime> m_rt;
};

}  // namespace iUtc
