// force_fp32.hip -- launchers of the FP32 pair kernels (plain and mixed steps,
// field-only evaluation).  The kernel itself is in force_tmpl.h.
#include "force_tmpl.h"

namespace utcnb {

namespace {
const Tile kTiles[] = {
    {1, 1, "k1s1"}, {2, 1, "k2s1"}, {4, 1, "k4s1"}, {1, 2, "k1s2"}, {2, 2, "k2s2"},
    {1, 4, "k1s4"}, {2, 4, "k2s4"}, {1, 8, "k1s8"}, {2, 8, "k2s8"}, {1, 16, "k1s16"},
};
}  // namespace

int tile_count() { return (int)(sizeof(kTiles) / sizeof(kTiles[0])); }
const Tile& tile_get(int id) { return kTiles[id]; }

// Enough waves to keep every SIMD busy: 256 CUs x 4 SIMDs x 2 waves.
int tile_default(int64_t n, int cu_count)
{
    const int64_t want = (int64_t)(cu_count > 0 ? cu_count : 256) * 8;
    const int64_t waves_k2 = (n + 127) / 128;
    if (waves_k2 >= want)
        return 1;  // k2s1
    const int64_t waves_k1 = (n + 63) / 64;
    if (waves_k1 >= want)
        return 0;  // k1s1
    if (waves_k1 * 4 >= want)
        return 5;  // k1s4
    if (waves_k1 * 8 >= want)
        return 7;  // k1s8
    return 9;      // k1s16
}

hipError_t launch_fp32_step(const StepArgs& a, const float4* pos_old, float4* pos_new,
                            float4* vel, hipStream_t st)
{
    EpiFP32 epi{pos_old, pos_new, vel, (float)a.dt, (float)a.damping, a.kick};
    if (a.newton)
        return launch_force<false, true>(a.tile, pos_old, a.n, a.n_pad, a.eps2f, epi, st);
    return launch_force<false, false>(a.tile, pos_old, a.n, a.n_pad, a.eps2f, epi, st);
}

hipError_t launch_mixed_step(const StepArgs& a, const float4* snap_old, float4* snap_new,
                             double4* pos, double4* vel, hipStream_t st)
{
    EpiMixed epi{pos, vel, snap_new, a.dt, a.damping, a.kick};
    if (a.newton)
        return launch_force<true, true>(a.tile, snap_old, a.n, a.n_pad, a.eps2f, epi, st);
    return launch_force<true, false>(a.tile, snap_old, a.n, a.n_pad, a.eps2f, epi, st);
}

hipError_t launch_fp32_accel(const StepArgs& a, const float4* src, bool compensated,
                             double* acc3, hipStream_t st)
{
    EpiAccel epi{acc3};
    if (compensated) {
        if (a.newton)
            return launch_force<true, true>(a.tile, src, a.n, a.n_pad, a.eps2f, epi, st);
        return launch_force<true, false>(a.tile, src, a.n, a.n_pad, a.eps2f, epi, st);
    }
    if (a.newton)
        return launch_force<false, true>(a.tile, src, a.n, a.n_pad, a.eps2f, epi, st);
    return launch_force<false, false>(a.tile, src, a.n, a.n_pad, a.eps2f, epi, st);
}

}  // namespace utcnb
