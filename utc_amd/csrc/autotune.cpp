// autotune.cpp -- picks the force-kernel tile (targets per lane K, source
// split S) for a problem size by timing every tile on the current device.
//
// Results are cached per (device, precision, size class) in memory and,
// optionally, in a text file with one line per entry:
//   <device-arch> <precision> <n> <tile-name> <ms-per-step>
// The size class is n rounded up to the source padding (1024 bodies).  Tiles
// are stored by name, so a file written by a build with another tile table
// maps onto this one; lines naming a tile this build lacks are ignored.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "internal.h"

namespace {

using Key = std::tuple<std::string, int, long long>;
std::mutex g_mu;
std::map<Key, std::pair<int, double>> g_cache;

std::string device_arch()
{
    int d = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&d) != hipSuccess || hipGetDeviceProperties(&prop, d) != hipSuccess)
        return "unknown";
    std::string a = prop.gcnArchName;
    size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

int tile_by_name(const char* name)
{
    for (int t = 0; t < utcnb::tile_count(); ++t)
        if (std::strcmp(utcnb::tile_get(t).name, name) == 0)
            return t + 1;
    return 0;
}

void load_file(const char* path)
{
    FILE* f = std::fopen(path, "r");
    if (!f)
        return;
    char arch[64], name[64];
    int prec;
    long long n;
    double ms;
    while (std::fscanf(f, "%63s %d %lld %63s %lf", arch, &prec, &n, name, &ms) == 5) {
        const int tile = tile_by_name(name);
        if (tile > 0)
            g_cache[Key(arch, prec, n)] = {tile, ms};
    }
    std::fclose(f);
}

void save_file(const char* path)
{
    std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f)
        return;
    for (auto& kv : g_cache)
        std::fprintf(f, "%s %d %lld %s %.17g\n", std::get<0>(kv.first).c_str(),
                     std::get<1>(kv.first), std::get<2>(kv.first),
                     utcnb::tile_get(kv.second.first - 1).name, kv.second.second);
    std::fclose(f);
    std::rename(tmp.c_str(), path);
}

}  // namespace

extern "C" int utcnb_autotune(int n, int precision, const char* cache_path, int* best_tile,
                              double* best_ms)
{
    using namespace utcnb;
    if (n < 1 || !best_tile)
        return fail(UTCNB_EINVAL, "utcnb_autotune: n < 1 or null output");
    if (precision != UTCNB_PREC_FP32 && precision != UTCNB_PREC_MIXED)
        return fail(UTCNB_EINVAL, "utcnb_autotune: tiles exist for FP32 and MIXED only");
    const long long cls = (long long)pad_j(n);
    const std::string arch = device_arch();
    std::lock_guard<std::mutex> lock(g_mu);
    if (cache_path)
        load_file(cache_path);
    auto it = g_cache.find(Key(arch, precision, cls));
    if (it != g_cache.end()) {
        *best_tile = it->second.first;
        if (best_ms)
            *best_ms = it->second.second;
        return UTCNB_OK;
    }
    // synthetic system of the class size: uniform cube, equal masses
    std::vector<double> pos(4 * (size_t)cls), vel(4 * (size_t)cls);
    int rc = utcnb_generate(UTCNB_GEN_CUBE, (int)cls, 12345, 0.01, 0, 0, pos.data(), vel.data(),
                            nullptr);
    if (rc)
        return fail(rc, "utcnb_autotune: generator failed");
    const int reps = cls >= 65536 ? 3 : (cls >= 8192 ? 10 : 50);
    int best = 0;
    double bms = 1e300;
    for (int t = 1; t <= tile_count(); ++t) {
        utcnb_config c{};
        c.n = (int)cls;
        c.precision = precision;
        c.kick = UTCNB_KICK_PHYSICAL;
        c.flags = UTCNB_FLAG_NO_NEWTON;
        c.tile = t;
        c.eps2 = 1e-4;
        c.damping = 1.0;
        c.dt = 1e-4;
        utcnb_system* s = nullptr;
        if ((rc = utcnb_create(&c, nullptr, &s))) {
            if (rc == UTCNB_EINVAL && (tile_get(t - 1).sym || tile_get(t - 1).cap > 0))
                continue;  // workspace too large, or too many bodies for the tile: not a candidate
            return rc;
        }
        double ms = 0.0;
        int launches = 0;
        rc = utcnb_set_state(s, pos.data(), vel.data(), UTCNB_F64, UTCNB_HOST);
        if (!rc)
            rc = utcnb_step(s, 1);
        if (!rc)
            rc = utcnb_step(s, reps);
        if (!rc)
            rc = utcnb_last_kernel_ms(s, &ms, &launches);
        utcnb_destroy(s);
        if (rc)
            return rc;
        ms /= launches;
        if (ms < bms) {
            bms = ms;
            best = t;
        }
    }
    if (best == 0)
        return fail(UTCNB_EHIP, "utcnb_autotune: no tile could be timed");
    g_cache[Key(arch, precision, cls)] = {best, bms};
    if (cache_path)
        save_file(cache_path);
    *best_tile = best;
    if (best_ms)
        *best_ms = bms;
    return UTCNB_OK;
}
