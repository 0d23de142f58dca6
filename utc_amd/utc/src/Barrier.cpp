// Barrier.cpp -- see Barrier.h.
#include "Barrier.h"

#include "TaskRuntime.h"

namespace iUtc {

void intra_Barrier()
{
    if (TaskRuntime* rt = currentRuntime())
        rt->barrier();
}

void inter_Barrier() { intra_Barrier(); }

}  // namespace iUtc
