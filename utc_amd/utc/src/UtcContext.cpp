// UtcContext.cpp -- see UtcContext.h.
#include "UtcContext.h"

#include <hip/hip_runtime_api.h>

namespace iUtc {

UtcContext& UtcContext::getContext()
{
    static UtcContext ctx;
    return ctx;
}

UtcContext& UtcContext::getContext(int, char**) { return getContext(); }

int UtcContext::getNumGpus() const
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

}  // namespace iUtc
