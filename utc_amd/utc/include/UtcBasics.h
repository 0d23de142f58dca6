// UtcBasics.h -- basic types of the UTC task surface (single-process build).
// Surface of the reference's include/UtcBasics.h (types only; the MPI and
// scoped-data switches of that header do not exist here).
#pragma once

namespace iUtc {
typedef int ProcRank_t;
typedef int ThreadRank_t;
typedef int TaskId_t;
}  // namespace iUtc
