// TaskBase.h -- task kinds and the untyped part of a task.
// Surface of include/TaskBase.h:22-147 (the parts a single process needs).
#pragma once

#include <string>
#include <vector>

#include "UtcBasics.h"

namespace iUtc {

enum class TaskType { unknown = 0, cpu_task, gpu_task };

class TaskBase {
public:
    virtual ~TaskBase() = default;
    std::string getName() const { return m_Name; }
    TaskId_t getTaskId() const { return m_TaskId; }
    TaskType getTaskType() const { return m_TaskType; }
    int getNumLocalThreads() const { return m_numLocalThreads; }
    int getNumTotalThreads() const { return m_numLocalThreads; }
    int getNumProcesses() const { return 1; }
    ProcRank_t getCurrentProcRank() const { return 0; }
    ProcRank_t getMainResideProcess() const { return 0; }

protected:
    std::string m_Name;
    TaskId_t m_TaskId = -1;
    TaskType m_TaskType = TaskType::unknown;
    int m_numLocalThreads = 0;
};

// The task the calling worker thread belongs to (nullptr on the main thread).
TaskBase* getCurrentTask();

}  // namespace iUtc
