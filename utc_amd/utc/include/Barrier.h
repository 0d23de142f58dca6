// Barrier.h -- synchronisation among the worker threads of one task.
// Surface of include/Barrier.h:66-69; intra_Barrier() must be called by every
// local thread of the current task (a no-op outside a task).
#pragma once

namespace iUtc {
void intra_Barrier();
void inter_Barrier();  // single process: same as intra_Barrier()
}  // namespace iUtc

using iUtc::intra_Barrier;
using iUtc::inter_Barrier;
