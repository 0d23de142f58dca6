// ProcList.h -- the list of process ranks a task runs on; one worker thread
// per entry that names this process.  Surface of include/ProcList.h:10-80.
#pragma once

#include <vector>

#include "UtcBasics.h"

namespace iUtc {

class ProcList {
public:
    ProcList();
    explicit ProcList(const std::vector<ProcRank_t>& ranks);
    ProcList(ProcRank_t proc);
    ProcList(int size, ProcRank_t proc);
    ProcList(int size, const int* ranks);
    ProcList(const ProcList&) = default;
    ProcList& operator=(const ProcList&) = default;

    void push_back(int rank);
    ProcRank_t getProc(int index) const;
    int getNumProcs() const;
    void getProcListVector(std::vector<ProcRank_t>& out) const;
    bool hasProc(int proc) const;
    void setProc(unsigned int index, unsigned int value);
    bool operator==(const ProcList& other) const;
    bool operator!=(const ProcList& other) const;

private:
    std::vector<ProcRank_t> m_procList;
};

}  // namespace iUtc
