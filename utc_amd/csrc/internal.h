// internal.h -- shared declarations of the gfx950 nbody library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "utcnb.h"

namespace utcnb {

// ---- error plumbing ------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define UTCNB_HIP(call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::utcnb::fail(UTCNB_EHIP, std::string(#call) + ": " +                 \
                                                 hipGetErrorString(e_));                 \
    } while (0)

// ---- layout constants ------------------------------------------------------
// Source (j) arrays of the fast kernels are padded to a multiple of
// kPadJ bodies with zero-mass bodies placed far away, so every wave streams a
// whole number of 64-body chunks and every j-split is chunk aligned.
constexpr int kChunk = 64;           // j per compensated partial sum
constexpr int kPadJ = 1024;          // = kChunk * largest j-split (16 waves)
constexpr float kPadCoord = 1.0e18f; // position of padding bodies
inline int64_t pad_j(int64_t n) { return (n + kPadJ - 1) / kPadJ * kPadJ; }

// ---- force-kernel tiles ---------------------------------------------------
// A tile is (K bodies i per lane, S waves splitting the j range per workgroup).
struct Tile {
    int k;       // targets per lane
    int s;       // source split (one-pass tiles) or waves per wave block (symmetric tiles)
    const char* name;
    bool sym;    // evaluates each unordered pair once (sym.hip)
    int rs = 1;  // symmetric tiles: rotation chunks per sub-unit (1, 2, 4, 8 or 16; K <= 4 when > 1)
    int tl = 0;  // small-system tiles (force_small.h): target lanes per wave; 0 for the other tiles
    int cap = 0; // small-system tiles: the most bodies (padded) their LDS holds
};
int tile_count();
const Tile& tile_get(int id);
// Heuristic default: enough waves to cover the chip (256 CUs x 8 waves).
// Device memory the symmetric kernels may take for their partial sums (force
// and VJP each): a quarter of the device, at most 64 GiB (16 GiB when the
// device size is unknown).  Larger systems use one-pass kernels.
size_t sym_partial_budget(size_t device_bytes);
int tile_default(int64_t n, int cu_count, size_t partial_budget);

// ---- launchers (force.hip) ------------------------------------------------
struct StepArgs {
    int n;          // bodies
    int n_pad;      // padded j length (fast kernels)
    float eps2f;
    double eps2;
    double dt, damping;
    int kick;       // UTCNB_KICK_*
    bool newton;
    int tile;       // tile id
    float* ipart;   // symmetric tiles: target-side partials [R][n_pad][3]
    float* jpart;   // symmetric tiles: reaction partials
    // symmetric tiles evaluated in chunks of wave blocks (0: one launch)
    int chunks;
    int chunk_r;    // waves per wave block in a chunk
    float* cipart;  // target-side partials [chunk_r][n_pad][3]
    float* cjpart;  // reaction partials of one chunk's wave blocks
    double* react;  // reactions summed per body over the chunks, FP64 [3n]
};
// Waves per wave block of a symmetric tile for a system of n_pad bodies.
int sym_tile_r(const Tile& t, int n_pad);
// Workspace of a symmetric tile (floats).
size_t sym_ipart_floats(int n_pad, int K, int R);
size_t sym_jpart_floats(int n_pad, int K, int rs = 1);

// Reference arithmetic, one fused step (pos_new/vel written), T = float/double.
hipError_t launch_ref_step_f32(const float4* pos_old, float4* pos_new, float4* vel, int n,
                               float dt, float eps2, float damping, int block,
                               hipStream_t st);
hipError_t launch_ref_step_f64(const double4* pos_old, double4* pos_new, double4* vel, int n,
                               double dt, double eps2, double damping, int block,
                               hipStream_t st);

// Plain FP32: forces from pos_old (padded float4), integrate into pos_new / vel.
hipError_t launch_fp32_step(const StepArgs& a, const float4* pos_old, float4* pos_new,
                            float4* vel, hipStream_t st);
// Mixed: forces from snap_old (padded float4), FP64 state updated in place,
// new snapshot written to snap_new.
hipError_t launch_mixed_step(const StepArgs& a, const float4* snap_old, float4* snap_new,
                             double4* pos, double4* vel, hipStream_t st);
// FP64: forces from pos_old (padded double4) into pos_new / vel.
hipError_t launch_fp64_step(const StepArgs& a, const double4* pos_old, double4* pos_new,
                            double4* vel, hipStream_t st);
// Accelerations only (FP64 out [3n]) for each arithmetic.
hipError_t launch_fp32_accel(const StepArgs& a, const float4* src, bool compensated,
                             double* acc3, hipStream_t st);
hipError_t launch_fp64_accel(const StepArgs& a, const double4* src, double* acc3,
                             hipStream_t st);
hipError_t launch_ref_accel_f32(const float4* src, int n, float eps2, double* acc3,
                                hipStream_t st);
hipError_t launch_ref_accel_f64(const double4* src, int n, double eps2, double* acc3,
                                hipStream_t st);

// ---- target slices: bodies [i0, i0 + nt) against all n sources ---------------
// Reference arithmetic: pos_new / vel are indexed by the slice (t = i - i0);
// the caller offsets full-size arrays.  acc3 rows are slice rows.
hipError_t launch_ref_step_slice_f32(const float4* pos_old, float4* pos_new, float4* vel, int n, int i0, int nt,
                                     float dt, float eps2, float damping, int block, hipStream_t st);
hipError_t launch_ref_step_slice_f64(const double4* pos_old, double4* pos_new, double4* vel, int n, int i0,
                                     int nt, double dt, double eps2, double damping, int block, hipStream_t st);
hipError_t launch_ref_accel_slice_f32(const float4* src, int n, int i0, int nt, float eps2, double* acc3,
                                      hipStream_t st);
hipError_t launch_ref_accel_slice_f64(const double4* src, int n, int i0, int nt, double eps2, double* acc3,
                                      hipStream_t st);
// Fast paths on one-pass tiles (`tile`; the symmetric ones do not split by
// target): full-size state arrays, updated for the slice only.
int slice_tile_default(int64_t nt, int cu_count);
hipError_t launch_fp32_step_slice(const StepArgs& a, int tile, int i0, int nt, const float4* pos_old,
                                  float4* pos_new, float4* vel, hipStream_t st);
hipError_t launch_mixed_step_slice(const StepArgs& a, int tile, int i0, int nt, const float4* snap_old,
                                   float4* snap_new, double4* pos, double4* vel, hipStream_t st);
hipError_t launch_fp32_accel_slice(const StepArgs& a, int tile, int i0, int nt, const float4* src,
                                   bool compensated, double* acc3, hipStream_t st);
// FP64: the source split of the whole system, so slices are bitwise the full field's rows.
hipError_t launch_fp64_step_slice(const StepArgs& a, int i0, int nt, const double4* pos_old, double4* pos_new,
                                  double4* vel, hipStream_t st);
hipError_t launch_fp64_accel_slice(const StepArgs& a, int i0, int nt, const double4* src, double* acc3,
                                   hipStream_t st);

// ---- one share of a symmetric-tile system sharded by wave blocks (sym.hip) -----
// Wave blocks [w0, w1) of share `share` of `shares` (the first NW % shares
// shares take one block more).
void sym_share_blocks(int NW, int shares, int share, int* w0, int* w1);
// Waves per wave block for a share of nw_share blocks (R of the tile or more).
int sym_share_splits(int nw_share, int R, int K, int NW);
// The share's pair work (target-side partials into ipart [Rs][n_pad][3],
// reactions into jpart) and its reactions summed per body into react3 (FP64 [3n]).
// jpart_share_rows: jpart holds the share's wave blocks' rows only (else all NW);
// accumulate: react3 += the share's reactions (else react3 =).
// compact_ipart: the share's target-side partials go to its own contiguous
// region of ipart (else to the whole system's row layout).
hipError_t launch_sym_share_pairs(const Tile& t, bool newton, const float4* src, int n, int n_pad, float eps2,
                                  int shares, int share, int Rs, float* ipart, bool compact_ipart, float* jpart,
                                  bool jpart_share_rows, double* react3, bool accumulate, hipStream_t st);
// Reaction partials of a chunk of nw wave blocks (floats).
inline size_t sym_chunk_jpart_floats(int nw, int NW, int K) { return (size_t)nw * (NW / 2) * 64 * K * 3; }
// Bodies [start, start + count): target-side partials + react_sum3 [3*count],
// then the MIXED epilogue (FP64 state in place, new snapshot into snap_new).
hipError_t launch_sym_share_finish_mixed(const StepArgs& a, int shares, int share, int Rs, int start, int count,
                                         const float* ipart, const double* react_sum3, float4* snap_new,
                                         double4* pos, double4* vel, hipStream_t st);

// ---- adjoint (adjoint.hip) ------------------------------------------------
// jrec: 2 records per body {x,y,z,m}{ax,ay,az,0}, n_pad bodies.
hipError_t launch_vjp_f32(const float4* jrec, int n, int n_pad, double eps2, double* xbar3,
                          double* mbar, int cu_count, hipStream_t st);
hipError_t launch_vjp_f64(const double4* jrec, int n, int n_pad, double eps2, double* xbar3,
                          double* mbar, int cu_count, hipStream_t st);
hipError_t launch_loss_seed(const double4* pos, const double* tgt3, int n, double* xbar3,
                            double* vbar3, double* mbar, double* part, double* loss,
                            hipStream_t st);
hipError_t launch_adjoint_pre(bool fp32_records, const double4* pos, const double* acc3, int n,
                              double dt, double damping, int kick, double* xbar3, double* vbar3,
                              double* mbar, void* jrec, hipStream_t st);
hipError_t launch_jrec_pad(bool fp32_records, void* jrec, int n, int n_pad, hipStream_t st);

// ---- symmetric adjoint pair sum (sym_vjp.hip): FP32 records, K = 4 ----------
bool sym_vjp_applies(int n_pad);                 // n_pad a multiple of 512
int sym_vjp_splits(int n_pad, int cu_count);     // waves per wave block
size_t sym_vjp_ipart_float4s(int n_pad, int R);
size_t sym_vjp_jpart_float4s(int n_pad);
// xbar3 / mbar (first n bodies) += the pair sums; ipart / jpart are workspace.
hipError_t launch_sym_vjp(const float4* jrec, int n, int n_pad, float eps2, int R, float4* ipart,
                          float4* jpart, double* xbar3, double* mbar, hipStream_t st);

// ---- batched systems (batch.hip) ---------------------------------------------
struct BatchGeom {
    const int* off;    // first body of each system (device)
    const int* n;      // bodies per system
    const int* chunks; // padded bodies / 64
    const int2* items; // (system, target tile) in workgroup order
    const int* start;  // items of workgroup b: [start[b], start[b+1])
};
hipError_t launch_batch_step(int W, const BatchGeom& g, int nblocks, const float4* snap_old,
                             float4* snap_new, double4* pos, double4* vel, float eps2, double dt,
                             double damping, int kick, hipStream_t st);
hipError_t launch_batch_vjp(int W, const BatchGeom& g, int nblocks, const float4* jrec, float eps2,
                            double* xbar3, double* mbar, hipStream_t st);
// Batched symmetric kernels (sym.hip, sym_vjp.hip): K = 4 targets per lane;
// every system starts on a multiple of kBatchSymAlign bodies and is padded to
// one, so it is a whole number of wave blocks.
constexpr int kBatchSymK = 4;
constexpr int kBatchSymAlign = 64 * kBatchSymK;
struct BatchSymGeom {
    const int4* items;        // (system, wave block, split, splits of the system)
    const int* start;         // items of wave g: [start[g], start[g+1])
    const int* off;           // first body of each system (padded layout)
    const int* n;             // bodies per system
    const int* nw;            // wave blocks per system
    const int* rs;            // splits per system
    const long long* ipo;     // force partials: first float of the system's target-side block
    const long long* jpo;     //   ... and of its reaction block
    const long long* vipo;    // VJP partials (float4 units), target side
    const long long* vjpo;    //   ... reactions
    const int* sys_of_block;  // system of each kBatchSymAlign-body block
    int* counter;             // work counter of the dynamic schedule, or null (static lists)
    int nitems;               // items in the dynamic schedule
    int body0;                // the reductions cover bodies [body0, body0 + total)
};
hipError_t launch_batch_counter_reset(int* counter, hipStream_t st);
hipError_t launch_batch_sym_step(const BatchSymGeom& g, int nwaves, int total, const float4* snap_old,
                                 float4* snap_new, double4* pos, double4* vel, float eps2, double dt,
                                 double damping, int kick, float* ipart, float* jpart, hipStream_t st);
hipError_t launch_batch_sym_vjp(const BatchSymGeom& g, int nwaves, int total, const float4* jrec,
                                float eps2, float4* ipart, float4* jpart, double* xbar3, double* mbar,
                                hipStream_t st);
hipError_t launch_batch_loss(const BatchGeom& g, int nsys, const double4* pos, const double* tgt3,
                             double* xbar3, double* vbar3, double* mbar, double* loss,
                             hipStream_t st);
hipError_t launch_batch_pad(const BatchGeom& g, int nsys, double4* pos, double4* vel, float4* s0,
                            float4* s1, hipStream_t st);

// ---- helpers (aux.hip) ----------------------------------------------------
// Fill padding bodies [n, n_pad) of a float4 / double4 source array.
hipError_t launch_pad_f32(float4* p, int n, int n_pad, hipStream_t st);
hipError_t launch_pad_f64(double4* p, int n, int n_pad, hipStream_t st);
// FP64 state -> FP32 snapshot (x,y,z,m rounded to nearest).
hipError_t launch_snapshot(const double4* pos, float4* snap, int n, hipStream_t st);
// Type conversions between state arrays.
hipError_t launch_f32_to_f64(const float4* src, double4* dst, int n, hipStream_t st);
hipError_t launch_f64_to_f32(const double4* src, float4* dst, int n, hipStream_t st);
// FP64 diagnostics: kinetic, potential (pairs i<j) and momentum, accumulated
// into out[5] = {K, W, Px, Py, Pz} (device, zeroed by the launcher).
hipError_t launch_energy(const double4* pos, const double4* vel, int n, double eps2,
                         double* out5, hipStream_t st);

}  // namespace utcnb
