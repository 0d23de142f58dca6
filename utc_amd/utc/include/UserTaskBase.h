// UserTaskBase.h -- base class of user tasks.
//
// Surface of the reference's include/UserTaskBase.h:28-104: a user task
// derives from UserTaskBase, is default constructible, and defines exactly one
// initImpl(...) and one runImpl(...) overload; Task<T>::init/run bind those.
// The thread-local members tell each worker thread who it is; GPU tasks also
// get the device and the HIP stream the worker thread owns.
#pragma once

#include <hip/hip_runtime_api.h>

class UserTaskBase {
public:
    UserTaskBase() = default;
    virtual ~UserTaskBase() = default;

    virtual void initImpl() {}
    virtual void runImpl() {}

    // Shared shape of the task; set once by the creating thread before any
    // worker starts.
    void setTaskShape(int numLocalThreads, int numProcesses, int numTotalProcesses,
                      int numTotalThreads);
    // Called on each worker thread before its first job / after its last one;
    // touches thread-local members only.
    void preInit(int lrank, int trank, int prank, hipStream_t stream, int device);
    void preExit();

    static thread_local int __localThreadId;
    static thread_local int __globalThreadId;
    static thread_local int __processIdInWorld;
    static thread_local int __processIdInGroup;
    int __numLocalThreads = 0;
    int __numGlobalThreads = 0;
    int __numWorldProcesses = 0;
    int __numGroupProcesses = 0;

    static thread_local hipStream_t __streamId;
    static thread_local int __deviceId;
};
