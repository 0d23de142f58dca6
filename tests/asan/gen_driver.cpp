// Host-only driver of the library's input generators (utc_amd/csrc/gen.cpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitizers.py.  Exact-size buffers, so any write past the
// documented extents (4n pos, 4n vel, 3n targets) is reported.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "utcnb.h"

namespace utcnb {
static std::string g_msg;
int fail(int code, const std::string& msg)  // gen.cpp's error hook (capi.cpp in the library)
{
    g_msg = msg;
    return code;
}
}  // namespace utcnb

static int check(bool ok, const char* what)
{
    if (!ok)
        std::printf("FAILED: %s\n", what);
    return ok ? 0 : 1;
}

int main()
{
    int bad = 0;
    const int sizes[] = {1, 2, 3, 17, 1000};
    for (int n : sizes) {
        std::vector<double> pos(4 * (size_t)n), vel(4 * (size_t)n), tgt(3 * (size_t)n);
        for (int kind : {UTCNB_GEN_CUBE, UTCNB_GEN_CUBE_LOGM, UTCNB_GEN_PLUMMER}) {
            const int rc = utcnb_generate(kind, n, 42 + n, 0.01, 0.5, 2.0, pos.data(), vel.data(), nullptr);
            bad += check(rc == UTCNB_OK, "generate");
            double msum = 0.0;
            for (size_t i = 0; i < pos.size(); ++i)
                bad += check(std::isfinite(pos[i]) && std::isfinite(vel[i]), "finite");
            for (int i = 0; i < n; ++i)
                msum += pos[4 * (size_t)i + 3];
            if (kind != UTCNB_GEN_CUBE_LOGM)
                bad += check(std::fabs(msum - 1.0) < 1e-12, "total mass 1");
        }
        bad += check(utcnb_generate(UTCNB_GEN_TARGETS, n, 7, 0.5, 0, 0, nullptr, nullptr, tgt.data()) == UTCNB_OK,
                     "targets");
    }
    std::vector<double> p(8), v(8), t(6);
    bad += check(utcnb_generate(UTCNB_GEN_CUBE, 0, 0, 0, 0, 0, p.data(), v.data(), nullptr) == UTCNB_EINVAL, "n = 0");
    bad += check(!utcnb::g_msg.empty(), "message set");
    bad += check(utcnb_generate(UTCNB_GEN_TARGETS, 2, 0, 1, 0, 0, nullptr, nullptr, nullptr) == UTCNB_EINVAL, "no out3");
    bad += check(utcnb_generate(UTCNB_GEN_PLUMMER, 2, 0, 0, 0, 0, p.data(), nullptr, nullptr) == UTCNB_EINVAL, "no vel");
    bad += check(utcnb_generate(UTCNB_GEN_CUBE_LOGM, 2, 0, 0.01, 2.0, 1.0, p.data(), v.data(), nullptr) == UTCNB_EINVAL,
                 "bad mass range");
    bad += check(utcnb_generate(99, 2, 0, 0, 0, 0, p.data(), v.data(), t.data()) == UTCNB_EINVAL, "unknown kind");
    uint64_t st[4];
    utcnb_rng_seed(1, st);
    uint64_t x = 0;
    for (int i = 0; i < 1000; ++i)
        x ^= utcnb_rng_next(st) ^ utcnb_splitmix64(&x);
    std::printf(bad ? "host sanitizer driver: %d failures\n" : "host sanitizer driver: ok (%d)\n", bad);
    return bad ? 1 : 0;
}
