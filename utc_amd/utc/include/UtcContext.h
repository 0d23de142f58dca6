// UtcContext.h -- process-level runtime context (single process, rank 0).
// Surface of include/UtcContext.h:17-94; the MPI world of the reference is a
// single process here, so numProcs() == 1 and Barrier() only orders threads.
#pragma once

namespace iUtc {

class UtcContext {
public:
    static UtcContext& getContext();
    static UtcContext& getContext(int argc, char** argv);
    int numProcs() const { return 1; }
    int getProcRank() const { return 0; }
    void Barrier() {}
    int getNumGpus() const;

private:
    UtcContext() = default;
};

}  // namespace iUtc
