// ProcList.cpp -- see ProcList.h.
#include "ProcList.h"

#include <algorithm>

namespace iUtc {

ProcList::ProcList() = default;
ProcList::ProcList(const std::vector<ProcRank_t>& ranks) : m_procList(ranks) {}
ProcList::ProcList(ProcRank_t proc) : m_procList(1, proc) {}
ProcList::ProcList(int size, ProcRank_t proc) : m_procList(size > 0 ? size : 0, proc) {}
ProcList::ProcList(int size, const int* ranks)
{
    for (int i = 0; i < size; ++i)
        m_procList.push_back(ranks[i]);
}
void ProcList::push_back(int rank) { m_procList.push_back(rank); }
ProcRank_t ProcList::getProc(int index) const { return m_procList.at((size_t)index); }
int ProcList::getNumProcs() const { return (int)m_procList.size(); }
void ProcList::getProcListVector(std::vector<ProcRank_t>& out) const { out = m_procList; }
bool ProcList::hasProc(int proc) const
{
    return std::find(m_procList.begin(), m_procList.end(), proc) != m_procList.end();
}
void ProcList::setProc(unsigned int index, unsigned int value) { m_procList.at(index) = (ProcRank_t)value; }
bool ProcList::operator==(const ProcList& o) const { return m_procList == o.m_procList; }
bool ProcList::operator!=(const ProcList& o) const { return !(*this == o); }

}  // namespace iUtc
