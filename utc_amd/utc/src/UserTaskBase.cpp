// UserTaskBase.cpp -- per-worker identity of a user task (see UserTaskBase.h).
#include "UserTaskBase.h"

thread_local int UserTaskBase::__localThreadId = -1;
thread_local int UserTaskBase::__globalThreadId = -1;
thread_local int UserTaskBase::__processIdInWorld = -1;
thread_local int UserTaskBase::__processIdInGroup = -1;
thread_local hipStream_t UserTaskBase::__streamId = nullptr;
thread_local int UserTaskBase::__deviceId = -1;

void UserTaskBase::setTaskShape(int numLocalThreads, int numProcesses, int numTotalProcesses,
                                int numTotalThreads)
{
    __numLocalThreads = numLocalThreads;
    __numGlobalThreads = numTotalThreads;
    __numWorldProcesses = numTotalProcesses;
    __numGroupProcesses = numProcesses;
}

void UserTaskBase::preInit(int lrank, int trank, int prank, hipStream_t stream, int device)
{
    __localThreadId = lrank;
    __globalThreadId = trank;
    __processIdInWorld = prank;
    __processIdInGroup = prank;
    __streamId = stream;
    __deviceId = device;
}

void UserTaskBase::preExit()
{
    __localThreadId = -1;
    __globalThreadId = -1;
    __streamId = nullptr;
    __deviceId = -1;
}
