// force_fp32.hip -- launchers of the FP32 pair kernels (plain and mixed steps,
// field-only evaluation).  The kernel itself is in force_tmpl.h.
#include <algorithm>

#include "force_tmpl.h"

namespace utcnb {

namespace {
const Tile kTiles[] = {
    {1, 1, "k1s1", false},  {2, 1, "k2s1", false},  {4, 1, "k4s1", false},
    {1, 2, "k1s2", false},  {2, 2, "k2s2", false},  {1, 4, "k1s4", false},
    {2, 4, "k2s4", false},  {1, 8, "k1s8", false},  {2, 8, "k2s8", false},
    {1, 16, "k1s16", false},
    // symmetric (sym.hip): K targets per lane, R waves per wave block
    {4, 4, "sym4r4", true}, {4, 8, "sym4r8", true}, {8, 4, "sym8r4", true},
    {8, 8, "sym8r8", true}, {8, 16, "sym8r16", true}, {4, 16, "sym4r16", true},
    {4, 32, "sym4r32", true}, {8, 32, "sym8r32", true}, {4, 64, "sym4r64", true},
    {4, 128, "sym4r128", true}, {2, 32, "sym2r32", true}, {2, 64, "sym2r64", true},
    {8, 64, "sym8r64", true}, {8, 128, "sym8r128", true},
    // symmetric with each sub-unit's rotation cut into RS chunks: symKrRxRS
    {2, 64, "sym2r64x2", true, 2}, {4, 64, "sym4r64x2", true, 2}, {2, 128, "sym2r128x4", true, 4},
    {4, 128, "sym4r128x4", true, 4},
    // one kernel per step for small systems (force_small.h): A targets per
    // lane, TL target lanes per wave, W waves, at most CAP bodies:
    // smAtTLwWnCAPk; k = A, s = W
    {1, 8, "sm1t4w8n2k", false, 1, 4, 2048}, {1, 8, "sm1t8w8n2k", false, 1, 8, 2048},
    {1, 16, "sm1t8w16n8k", false, 1, 8, 8192}, {1, 16, "sm1t16w16n4k", false, 1, 16, 4096},
    {1, 16, "sm1t16w16n8k", false, 1, 16, 8192}, {2, 16, "sm2t16w16n8k", false, 1, 16, 8192},
    // more one-pass tiles (appended: ids are stable)
    {2, 16, "k2s16", false}, {4, 8, "k4s8", false}, {4, 16, "k4s16", false},
};
}  // namespace

int tile_count() { return (int)(sizeof(kTiles) / sizeof(kTiles[0])); }
const Tile& tile_get(int id) { return kTiles[id]; }

// Default tile, from sweeps of every tile on MI355X (mixed precision, time per
// step, best three; Plummer spheres, equal masses; log-uniform masses rank
// alike):
//   N = 262144: sym8r64 12.29, sym8r48 12.34, sym8r32 12.34 ms  (k2s8, one-pass: 20.7)
//   N = 131072: sym8r48 3.139, sym8r64 3.143, sym8r32 3.160 ms
//   N =  65536: sym8r64 0.830, sym8r48 0.830, sym8r32 0.871 ms
//   N =  32768: sym8r64 0.235, sym8r48 0.239, sym8r32 0.240 ms
//   N =  24576: sym8r64 148.0, sym4r32 150.9, sym4r64 152.5 us
//   N =  16384: sym4r64x2 75.9, sym4r64 77.5, sym4r128x2 79.3 us
//   N =  12288: sym2r32 51.8, sym4r64x2 51.8, sym4r128x2 52.4 us
//   N =   8192: sm2t16w16n8k 28.4, sm1t16w16n8k 29.0, sm4t8w16n8k 29.5 us  (sym2r32 29.7)
//   N =   6144: sm2t16w16n8k 21.4, sm1t8w16n8k 21.8, sm1t16w16n8k 22.4 us  (sym2r64x2 23.2)
//   N =   4096: sm1t16w16n4k 9.57, sm1t16w16n8k 9.59, sm2t8w16n8k 10.2 us  (sym2r32 16.0)
//   N =   3072: sm1t16w16n4k 7.88, sm1t16w16n8k 7.91, sm2t8w16n8k 8.42 us  (sym2r32 12.1)
//   N =   2048: sm1t8w16n8k 5.11, sm1t8w16n4k 5.44, sm1t8w8n2k 5.48 us  (sym2r32 10.9)
//   N =   1024: sm1t4w8n2k 3.94, sm1t8w8n2k 4.01, sm1t8w16n8k 4.17 us  (k1s16 10.1)
// (sym8r48, sym8r24 and several other small-system shapes were swept and
// are not in the table.)  Up to 8192 bodies the small-system kernels win:
// one launch per step against two, and their one-pass pair sums fill the
// chip where the symmetric kernels' 64-step rotations leave it waiting.  The
// symmetric tiles need enough wave blocks to fill the chip (K = 8 from 96
// bodies per CU, K = 4 from 64), and their partial sums must fit the
// device's budget (sym_partial_budget); above it the best one-pass tile
// (k2s8) is used.
size_t sym_partial_budget(size_t device_bytes)
{
    if (device_bytes == 0)
        return (size_t)16 << 30;
    return std::min<size_t>((size_t)64 << 30, device_bytes / 4);
}

int tile_default(int64_t n, int cu_count, size_t partial_budget)
{
    const int64_t cus = cu_count > 0 ? cu_count : 256;
    const int64_t n_pad = pad_j(n);
    const auto fits = [&](int k) {
        return n_pad <= (1 << 30) && n_pad % (128 * k) == 0 &&
               (sym_jpart_floats((int)n_pad, k) + sym_ipart_floats((int)n_pad, k, 64)) * sizeof(float) <=
                   partial_budget;
    };
    // from the sweeps above: each choice within 2% of the fastest tile at its size
    if (n <= 1024)
        return 28;  // sm1t4w8n2k
    if (n <= 2048)
        return 30;  // sm1t8w16n8k
    if (n <= 4096)
        return 31;  // sm1t16w16n4k
    if (n <= 8192)
        return 33;  // sm2t16w16n8k
    if (n >= 96 * cus && fits(8))
        return 22;  // sym8r64
    if (n >= 64 * cus && fits(4))
        return 25;  // sym4r64x2
    if (n >= 8 * cus && fits(2))
        return 20;  // sym2r32
    if ((n + 127) / 128 * 8 >= cus * 16)
        return 8;   // k2s8
    return 9;       // k1s16
}

template <class Epi>
hipError_t launch_sym(const Tile& t, bool newton, bool comp, const float4* src, int n, int n_pad, float eps2,
                      float* ipart, float* jpart, const Epi& epi, hipStream_t st);
template <class Epi>
hipError_t launch_sym_chunked(const Tile& t, bool newton, const float4* src, int n, int n_pad, float eps2,
                              int chunks, int Rs, float* ipart, float* jchunk, double* react3, const Epi& epi,
                              hipStream_t st);

hipError_t launch_fp32_step(const StepArgs& a, const float4* pos_old, float4* pos_new,
                            float4* vel, hipStream_t st)
{
    EpiFP32 epi{pos_old, pos_new, vel, (float)a.dt, (float)a.damping, a.kick};
    if (tile_get(a.tile).sym)
        return launch_sym(tile_get(a.tile), a.newton, false, pos_old, a.n, a.n_pad, a.eps2f, a.ipart,
                          a.jpart, epi, st);
    if (a.newton)
        return launch_force<false, true>(a.tile, pos_old, a.n, a.n_pad, a.eps2f, epi, st);
    return launch_force<false, false>(a.tile, pos_old, a.n, a.n_pad, a.eps2f, epi, st);
}

hipError_t launch_mixed_step(const StepArgs& a, const float4* snap_old, float4* snap_new,
                             double4* pos, double4* vel, hipStream_t st)
{
    EpiMixed epi{pos, vel, snap_new, a.dt, a.damping, a.kick};
    if (tile_get(a.tile).sym && a.chunks > 0)
        return launch_sym_chunked(tile_get(a.tile), a.newton, snap_old, a.n, a.n_pad, a.eps2f, a.chunks, a.chunk_r,
                                  a.cipart, a.cjpart, a.react, epi, st);
    if (tile_get(a.tile).sym)
        return launch_sym(tile_get(a.tile), a.newton, true, snap_old, a.n, a.n_pad, a.eps2f, a.ipart,
                          a.jpart, epi, st);
    if (a.newton)
        return launch_force<true, true>(a.tile, snap_old, a.n, a.n_pad, a.eps2f, epi, st);
    return launch_force<true, false>(a.tile, snap_old, a.n, a.n_pad, a.eps2f, epi, st);
}

// One-pass tile for a slice of nt targets against all sources (the
// symmetric tiles evaluate unordered pairs of the whole system and do not
// split by target).  From a sweep on MI355X, slices of N = 262144, ms per
// step (best three):
//   nt =   8192: k1s16 1.81, k1s8 1.98, k2s16 2.86
//   nt =  16384: k1s16 1.81, k1s8 2.12, k2s16 2.86
//   nt =  32768: k2s16 2.89, k2s8 3.22, k1s16 3.59
//   nt =  65536: k2s8 5.73, k2s16 5.84, k4s16 5.96   (k1s16 7.19)
//   nt = 131072: k2s8 10.76, k2s16 11.74, k4s16 11.89 (k1s16 14.59)
//   nt = 262144: k2s8 20.72, k2s16 23.65, k4s16 23.76
// The fewer the targets per CU, the more source splits keep the waves busy.
int slice_tile_default(int64_t nt, int cu_count)
{
    const int64_t cus = cu_count > 0 ? cu_count : 256;
    if (nt <= 64 * cus)
        return 9;   // k1s16
    if (nt <= 128 * cus)
        return 34;  // k2s16
    return 8;       // k2s8
}

hipError_t launch_fp32_step_slice(const StepArgs& a, int tile, int i0, int nt, const float4* pos_old,
                                  float4* pos_new, float4* vel, hipStream_t st)
{
    if (tile_get(tile).sym)
        return hipErrorInvalidValue;
    EpiFP32 epi{pos_old, pos_new, vel, (float)a.dt, (float)a.damping, a.kick};
    if (a.newton)
        return launch_force_range<false, true>(tile, pos_old, i0, nt, a.n_pad, a.eps2f, epi, st);
    return launch_force_range<false, false>(tile, pos_old, i0, nt, a.n_pad, a.eps2f, epi, st);
}

hipError_t launch_mixed_step_slice(const StepArgs& a, int tile, int i0, int nt, const float4* snap_old,
                                   float4* snap_new, double4* pos, double4* vel, hipStream_t st)
{
    if (tile_get(tile).sym)
        return hipErrorInvalidValue;
    EpiMixed epi{pos, vel, snap_new, a.dt, a.damping, a.kick};
    if (a.newton)
        return launch_force_range<true, true>(tile, snap_old, i0, nt, a.n_pad, a.eps2f, epi, st);
    return launch_force_range<true, false>(tile, snap_old, i0, nt, a.n_pad, a.eps2f, epi, st);
}

hipError_t launch_fp32_accel_slice(const StepArgs& a, int tile, int i0, int nt, const float4* src,
                                   bool compensated, double* acc3, hipStream_t st)
{
    if (tile_get(tile).sym)
        return hipErrorInvalidValue;
    EpiAccel epi{acc3, i0};
    if (compensated)
        return a.newton ? launch_force_range<true, true>(tile, src, i0, nt, a.n_pad, a.eps2f, epi, st)
                        : launch_force_range<true, false>(tile, src, i0, nt, a.n_pad, a.eps2f, epi, st);
    return a.newton ? launch_force_range<false, true>(tile, src, i0, nt, a.n_pad, a.eps2f, epi, st)
                    : launch_force_range<false, false>(tile, src, i0, nt, a.n_pad, a.eps2f, epi, st);
}

hipError_t launch_fp32_accel(const StepArgs& a, const float4* src, bool compensated,
                             double* acc3, hipStream_t st)
{
    EpiAccel epi{acc3};
    if (tile_get(a.tile).sym && a.chunks > 0) {
        if (!compensated)
            return hipErrorInvalidValue;  // the chunked step sums its partials in FP64
        return launch_sym_chunked(tile_get(a.tile), a.newton, src, a.n, a.n_pad, a.eps2f, a.chunks, a.chunk_r,
                                  a.cipart, a.cjpart, a.react, epi, st);
    }
    if (tile_get(a.tile).sym)
        return launch_sym(tile_get(a.tile), a.newton, compensated, src, a.n, a.n_pad, a.eps2f, a.ipart,
                          a.jpart, epi, st);
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
