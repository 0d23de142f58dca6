// batch_api.cpp -- the batched-systems part of the C ABI (utcnb_batch_*).
//
// The reference's analog is its multi-task driver: T independent
// Task<BodySystemSGPU> objects, one worker thread and stream each
// (benchapps/nbody/utc/nbody_mtask.cc:185-231).  Here the systems share one
// concatenated device state and one persistent launch per step (batch.hip);
// the sequence of launches of a run is captured once in a HIP graph and
// replayed.  Mode 1 keeps the per-system structure (one launch of the
// single-system kernels per system and step, spread over four streams like
// the reference's stream per task) as the baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <new>
#include <queue>
#include <string>
#include <vector>

#include "internal.h"

using namespace utcnb;

namespace {

constexpr int kWs[3] = {4, 8, 16};  // waves per work item of the one-pass kernels (batch tiles 1..3)
// Batch tiles 4 and 5: the symmetric kernels (sym.hip, sym_vjp.hip), work items
// (system, wave block, split) in a longest-first list, taken by persistent
// waves through an atomic work counter (4), or one wave per item for the
// hardware dispatcher to place (5).
constexpr int kSymPersist = 3, kSymFlat = 4, kSels = 5;
// Work units (K source blocks of 64) per symmetric item.  With the two
// concurrent system groups, 3 ran the 256-system graph 1.2% faster forward
// and 2.5% faster forward + adjoint than 2 (MI355X, tile 5; 1 and 4 in
// between).
constexpr int kSymUnitsPerItem = 3;
// The per-system baseline (mode 1) spreads its launches over this many streams,
// as the reference runs one task with its own stream per system; 4 is the
// number of hardware queues a process gets by default.
constexpr int kAuxStreams = 4;
// The symmetric schedules split the systems into this many contiguous groups
// of about equal pair count, each run on a stream of its own inside the
// captured graph, so that one group's work fills the tail of the other's
// steps (a step ends with the last long items, then the reduction).  On
// MI355X, 256 systems of 256..8192 bodies: two groups ran the 20-step graph
// 6.6% faster forward and 7.7% faster forward + adjoint than one; three or
// four groups were within 1.3% of two.
constexpr int kGroups = 2;

template <typename T>
int dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0)
        return UTCNB_OK;
    if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return fail(UTCNB_ENOMEM, "utcnb_batch: hipMalloc failed");
    }
    return UTCNB_OK;
}

template <typename T>
void dfree(T*& p)
{
    if (p)
        (void)hipFree((void*)p);
    p = nullptr;
}

}  // namespace

struct utcnb_batch {
    utcnb_config cfg{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int cu_count = 256;
    size_t partial_budget = (size_t)16 << 30;  // sym_partial_budget of the device
    int nsys = 0;
    std::vector<int> n, npad, off;
    long long total = 0, total_real = 0;
    double4* pos = nullptr;
    double4* vel = nullptr;
    float4* snap[2] = {nullptr, nullptr};
    int* d_off = nullptr;
    int* d_n = nullptr;
    int* d_chunks = nullptr;
    int2* d_items[3] = {nullptr, nullptr, nullptr};
    int* d_start[3] = {nullptr, nullptr, nullptr};
    int nblocks[3] = {0, 0, 0};
    int wsel = -1;  // kernel selection 0..kSels-1 (batch tile - 1), -1 = not tuned yet
    int align = kChunk;  // systems start on multiples of this (kBatchSymAlign for the symmetric kernels)
    // symmetric kernels: per-system geometry and partial sums
    int* d_nw = nullptr;
    int* d_rs = nullptr;
    long long* d_ipo = nullptr;
    long long* d_jpo = nullptr;
    long long* d_vipo = nullptr;
    long long* d_vjpo = nullptr;
    int* d_sysblk = nullptr;
    int4* d_sitems[2] = {nullptr, nullptr};
    int* d_sstart[2] = {nullptr, nullptr};
    int ngroups = 1;                   // system groups (up to kGroups, at most one per system)
    int gsys[kGroups + 1] = {};        // group g: systems [gsys[g], gsys[g+1])
    int gbody[kGroups + 1] = {};       //   and bodies [gbody[g], gbody[g+1])
    int sibase[2][kGroups] = {};       // first item of group g in d_sitems[k]
    int ssbase[2][kGroups] = {};       // first entry of group g in d_sstart[k]
    int snwaves[2][kGroups] = {};      // waves per launch of group g
    int snitems[kGroups] = {};         // items of group g
    int* d_counter = nullptr;  // [kGroups][2]: work counters of the persistent step and VJP kernels
    hipStream_t gstream[kGroups] = {};  // group g > 0 runs on gstream[g] (group 0 on stream)
    hipEvent_t gfork = nullptr, gjoin[kGroups] = {};
    float* sipart = nullptr;
    float* sjpart = nullptr;
    size_t svi_count = 0, svj_count = 0;  // VJP partials (float4), allocated with the adjoint workspace
    float4* svipart = nullptr;
    float4* svjpart = nullptr;
    bool has_state = false;
    // adjoint workspace
    int nck = 0, seg = 0;
    double4* ckpt = nullptr;
    double4* segpos = nullptr;
    double4* wpos = nullptr;
    double4* wvel = nullptr;
    float4* wsnap[2] = {nullptr, nullptr};
    float4* jrec = nullptr;
    double *xbar = nullptr, *vbar = nullptr, *mbar = nullptr, *tgt = nullptr, *loss = nullptr;
    // captured graph
    hipGraphExec_t exec = nullptr;
    int g_nsteps = -1, g_ckpt = -1, g_w = -1;
    bool g_adj = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t aux[kAuxStreams] = {};  // per-system baseline streams (created on first use)
    hipEvent_t evfork = nullptr, evjoin[kAuxStreams] = {};

    BatchGeom geom(int wi) const
    {
        return BatchGeom{d_off, d_n, d_chunks, d_items[wi], d_start[wi]};
    }
    // k: 0 persistent waves (dynamic), 1 one wave per item; which: 0 step, 1 VJP counter; g: group
    BatchSymGeom sym_geom(int k, int which, int g) const
    {
        return BatchSymGeom{d_sitems[k] + sibase[k][g], d_sstart[k] + ssbase[k][g], d_off, d_n, d_nw, d_rs,
                            d_ipo, d_jpo, d_vipo, d_vjpo, d_sysblk,
                            k == 0 ? d_counter + 2 * g + which : nullptr, snitems[g], gbody[g]};
    }
    int group_bodies(int g) const { return gbody[g + 1] - gbody[g]; }
    bool sym_ok = false;  // the symmetric kernels' geometry and force partials are set up
    bool has_sym() const { return sym_ok; }
};

namespace {

void release_adj(utcnb_batch* b)
{
    // a captured graph holds the addresses of the buffers freed below: drop
    // it, so the next mode 0 run captures again instead of replaying launches
    // on freed memory
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->g_nsteps = -1;
    dfree(b->ckpt);
    dfree(b->segpos);
    dfree(b->wpos);
    dfree(b->wvel);
    dfree(b->wsnap[0]);
    dfree(b->wsnap[1]);
    dfree(b->jrec);
    dfree(b->xbar);
    dfree(b->vbar);
    dfree(b->mbar);
    dfree(b->tgt);
    dfree(b->loss);
    dfree(b->svipart);
    dfree(b->svjpart);
    b->nck = b->seg = 0;
}

void release(utcnb_batch* b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    release_adj(b);
    dfree(b->pos);
    dfree(b->vel);
    dfree(b->snap[0]);
    dfree(b->snap[1]);
    dfree(b->d_off);
    dfree(b->d_n);
    dfree(b->d_chunks);
    for (int w = 0; w < 3; ++w) {
        dfree(b->d_items[w]);
        dfree(b->d_start[w]);
    }
    dfree(b->d_nw);
    dfree(b->d_rs);
    dfree(b->d_ipo);
    dfree(b->d_jpo);
    dfree(b->d_vipo);
    dfree(b->d_vjpo);
    dfree(b->d_sysblk);
    for (int k = 0; k < 2; ++k) {
        dfree(b->d_sitems[k]);
        dfree(b->d_sstart[k]);
    }
    dfree(b->sipart);
    dfree(b->sjpart);
    dfree(b->d_counter);
    for (int g = 0; g < kGroups; ++g) {
        if (b->gstream[g])
            (void)hipStreamDestroy(b->gstream[g]);
        if (b->gjoin[g])
            (void)hipEventDestroy(b->gjoin[g]);
        b->gstream[g] = nullptr;
        b->gjoin[g] = nullptr;
    }
    if (b->gfork)
        (void)hipEventDestroy(b->gfork);
    b->gfork = nullptr;
    if (b->ev0)
        (void)hipEventDestroy(b->ev0);
    if (b->ev1)
        (void)hipEventDestroy(b->ev1);
    for (int k = 0; k < kAuxStreams; ++k) {
        if (b->aux[k])
            (void)hipStreamDestroy(b->aux[k]);
        if (b->evjoin[k])
            (void)hipEventDestroy(b->evjoin[k]);
        b->aux[k] = nullptr;
        b->evjoin[k] = nullptr;
    }
    if (b->evfork)
        (void)hipEventDestroy(b->evfork);
    b->evfork = nullptr;
    if (b->own_stream && b->stream)
        (void)hipStreamDestroy(b->stream);
}

// Longest-processing-time schedule of (system, tile) items over `blocks`
// workgroups for W waves per item.  Item cost: ceil(chunks / W) source chunks
// per wave plus one for the reduction and epilogue.
void schedule(const utcnb_batch* b, int W, int blocks, std::vector<int2>& items,
              std::vector<int>& start)
{
    struct It {
        long long cost;
        int s, t;
    };
    std::vector<It> all;
    for (int s = 0; s < b->nsys; ++s) {
        const int C = b->npad[s] / kChunk;
        const long long cost = (C + W - 1) / W + 1;
        for (int t = 0; t < (b->n[s] + 63) / 64; ++t)
            all.push_back({cost, s, t});
    }
    std::stable_sort(all.begin(), all.end(), [](const It& a, const It& c) { return a.cost > c.cost; });
    using Load = std::pair<long long, int>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int k = 0; k < blocks; ++k)
        heap.push({0, k});
    std::vector<std::vector<int2>> per(blocks);
    for (const It& it : all) {
        Load l = heap.top();
        heap.pop();
        per[l.second].push_back(make_int2(it.s, it.t));
        heap.push({l.first + it.cost, l.second});
    }
    items.clear();
    start.assign(1, 0);
    for (int k = 0; k < blocks; ++k) {
        items.insert(items.end(), per[k].begin(), per[k].end());
        start.push_back((int)items.size());
    }
}

// Symmetric kernels: per system NW = npad / kBatchSymAlign wave blocks, each
// with K (NW/2 + 1) (unit, source block) sub-units split over R waves (about
// kSymUnitsPerItem units each); an item's cost is its number of sub-units.
// Items (system, wave block, split, R) are listed longest first; persistent
// waves take them in that order through an atomic work counter, or each item
// gets a wave of its own.
struct SymPlan {
    std::vector<int> nw, rs, sysblk;
    std::vector<long long> ipo, jpo, vipo, vjpo;
    long long ifloats = 0, jfloats = 0, vi4 = 0, vj4 = 0;
    std::vector<int4> items[2];  // groups one after another
    std::vector<int> start[2];   // per group, relative to the group's first item
    int ibase[2][kGroups] = {}, sbase[2][kGroups] = {}, nwaves[2][kGroups] = {}, nitems[kGroups] = {};
};

void plan_sym(const utcnb_batch* b, int persistent_waves, SymPlan& P)
{
    struct It {
        int cost;
        int4 item;
    };
    std::vector<It> all[kGroups];
    for (int s = 0; s < b->nsys; ++s) {
        int grp = 0;
        while (grp + 1 < b->ngroups && s >= b->gsys[grp + 1])
            ++grp;
        const int NW = b->npad[s] / kBatchSymAlign;
        const int sub = (NW / 2 + 1) * kBatchSymK;  // (unit, source block) sub-units per wave block
        const int per = kSymUnitsPerItem * kBatchSymK;
        const int R = std::max(1, (sub + per - 1) / per);
        P.nw.push_back(NW);
        P.rs.push_back(R);
        P.ipo.push_back(P.ifloats);
        P.jpo.push_back(P.jfloats);
        P.vipo.push_back(P.vi4);
        P.vjpo.push_back(P.vj4);
        P.ifloats += (long long)R * b->npad[s] * 3;
        P.jfloats += (long long)NW * (NW / 2) * kBatchSymAlign * 3;
        P.vi4 += (long long)R * b->npad[s];
        P.vj4 += (long long)NW * (NW / 2) * kBatchSymAlign;
        for (int blk = 0; blk < NW; ++blk)
            P.sysblk.push_back(s);
        for (int w = 0; w < NW; ++w)
            for (int r = 0; r < R; ++r) {
                const int q0 = (int)((long long)sub * r / R), q1 = (int)((long long)sub * (r + 1) / R);
                all[grp].push_back({q1 - q0, make_int4(s, w, r, R)});
            }
    }
    for (int k = 0; k < 2; ++k) {
        P.items[k].clear();
        P.start[k].clear();
    }
    for (int g = 0; g < b->ngroups; ++g) {
        std::vector<It>& L = all[g];
        std::stable_sort(L.begin(), L.end(), [](const It& a, const It& c) { return a.cost > c.cost; });
        const int cnt = (int)L.size();
        for (int k = 0; k < 2; ++k) {
            P.ibase[k][g] = (int)P.items[k].size();
            P.sbase[k][g] = (int)P.start[k].size();
            for (const It& it : L)
                P.items[k].push_back(it.item);
        }
        // one wave per item (start[w] = w)
        for (int w = 0; w <= cnt; ++w)
            P.start[1].push_back(w);
        P.nwaves[1][g] = cnt;
        // persistent waves take the same longest-first list through a work counter
        P.start[0].push_back(0);
        P.start[0].push_back(cnt);
        P.nwaves[0][g] = std::max(1, std::min(cnt, persistent_waves));
        P.nitems[g] = cnt;
    }
}

// Splits the systems into ngroups = min(kGroups, nsys) contiguous groups of
// about equal pair count: group g ends at the first system where the running
// sum of n_pad^2 reaches (g+1)/ngroups of the total, or just before it,
// whichever is nearer, keeping every group non-empty.
void split_groups(utcnb_batch* b)
{
    b->ngroups = std::min(kGroups, b->nsys);
    std::vector<double> pre(b->nsys + 1, 0.0);
    for (int s = 0; s < b->nsys; ++s)
        pre[s + 1] = pre[s] + (double)b->npad[s] * b->npad[s];
    b->gsys[0] = 0;
    for (int g = 1; g < b->ngroups; ++g) {
        const double want = pre[b->nsys] * g / b->ngroups;
        int m = b->gsys[g - 1] + 1;  // at least one system in group g - 1
        while (m < b->nsys && pre[m] < want)
            ++m;
        if (m > b->gsys[g - 1] + 1 && want - pre[m - 1] < pre[m] - want)
            --m;
        m = std::min(m, b->nsys - (b->ngroups - g));  // and one for each later group
        b->gsys[g] = m;
    }
    b->gsys[b->ngroups] = b->nsys;
    for (int g = 0; g <= b->ngroups; ++g)
        b->gbody[g] = g < b->ngroups ? b->off[b->gsys[g]] : (int)b->total;
}

template <typename T>
hipError_t upload(T** d, const std::vector<T>& h)
{
    if (dalloc(d, h.size()))
        return hipErrorOutOfMemory;
    return hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}

// Sets up the symmetric kernels; returns hipSuccess with b->sym_ok false when
// their partial sums (force or VJP) would exceed the device's
// sym_partial_budget; such batches use the one-pass kernels.
hipError_t setup_sym(utcnb_batch* b)
{
    split_groups(b);
    SymPlan P;
    plan_sym(b, b->cu_count * 16, P);
    const size_t fbytes = (size_t)(P.ifloats + P.jfloats) * sizeof(float);
    const size_t vbytes = (size_t)(P.vi4 + P.vj4) * sizeof(float4);
    if (fbytes > b->partial_budget || vbytes > b->partial_budget)
        return hipSuccess;
    hipError_t e = upload(&b->d_nw, P.nw);
    if (e == hipSuccess) e = upload(&b->d_rs, P.rs);
    if (e == hipSuccess) e = upload(&b->d_ipo, P.ipo);
    if (e == hipSuccess) e = upload(&b->d_jpo, P.jpo);
    if (e == hipSuccess) e = upload(&b->d_vipo, P.vipo);
    if (e == hipSuccess) e = upload(&b->d_vjpo, P.vjpo);
    if (e == hipSuccess) e = upload(&b->d_sysblk, P.sysblk);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = upload(&b->d_sitems[k], P.items[k]);
        if (e == hipSuccess)
            e = upload(&b->d_sstart[k], P.start[k]);
    }
    for (int g = 0; g < b->ngroups; ++g) {
        for (int k = 0; k < 2; ++k) {
            b->sibase[k][g] = P.ibase[k][g];
            b->ssbase[k][g] = P.sbase[k][g];
            b->snwaves[k][g] = P.nwaves[k][g];
        }
        b->snitems[g] = P.nitems[g];
    }
    if (e == hipSuccess && dalloc(&b->d_counter, 2 * kGroups))
        e = hipErrorOutOfMemory;
    if (e == hipSuccess)
        e = hipMemset(b->d_counter, 0, 2 * kGroups * sizeof(int));
    if (e == hipSuccess && b->ngroups > 1)
        e = hipEventCreateWithFlags(&b->gfork, hipEventDisableTiming);
    for (int g = 1; g < b->ngroups && e == hipSuccess; ++g) {
        e = hipStreamCreateWithFlags(&b->gstream[g], hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&b->gjoin[g], hipEventDisableTiming);
    }
    if (e == hipSuccess && (dalloc(&b->sipart, (size_t)P.ifloats) || dalloc(&b->sjpart, (size_t)P.jfloats)))
        e = hipErrorOutOfMemory;
    b->svi_count = (size_t)P.vi4;
    b->svj_count = (size_t)P.vj4;
    b->sym_ok = e == hipSuccess;
    return e;
}

int adj_workspace(utcnb_batch* b, int nck, int seg)
{
    if (b->xbar && b->nck >= nck && b->seg >= seg)
        return UTCNB_OK;
    release_adj(b);
    const size_t T = (size_t)b->total;
    int rc = dalloc(&b->ckpt, 2 * T * (size_t)nck) || dalloc(&b->segpos, T * (size_t)seg) ||
             dalloc(&b->wpos, T) || dalloc(&b->wvel, T) || dalloc(&b->wsnap[0], T) ||
             dalloc(&b->wsnap[1], T) || dalloc(&b->jrec, 2 * T) || dalloc(&b->xbar, 3 * T) ||
             dalloc(&b->vbar, 3 * T) || dalloc(&b->mbar, T) || dalloc(&b->tgt, 3 * T) ||
             dalloc(&b->loss, (size_t)b->nsys);
    if (!rc && b->has_sym())
        rc = dalloc(&b->svipart, b->svi_count) || dalloc(&b->svjpart, b->svj_count);
    if (rc) {
        release_adj(b);
        return fail(UTCNB_ENOMEM, "utcnb_batch_run: adjoint workspace allocation failed");
    }
    b->nck = nck;
    b->seg = seg;
    // the step kernels write real bodies only: pad the work state once
    hipError_t e = launch_batch_pad(b->geom(0), b->nsys, b->wpos, b->wvel, b->wsnap[0], b->wsnap[1],
                                    b->stream);
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_batch_run: padding: ") + hipGetErrorString(e));
    return UTCNB_OK;
}

// Tile id (single-system table) of K = 1 with the largest source split <= 16
// that keeps whole 64-body chunks per wave.
int per_system_tile(int npad)
{
    const int C = npad / kChunk;
    static const int ids[5] = {0, 3, 5, 7, 9};  // k1s1 k1s2 k1s4 k1s8 k1s16
    int best = 0;
    for (int e = 1; e < 5; ++e)
        if (C % (1 << e) == 0)
            best = e;
    return ids[best];
}

hipError_t aux_streams(utcnb_batch* b)
{
    if (b->aux[0])
        return hipSuccess;
    hipError_t e = hipEventCreateWithFlags(&b->evfork, hipEventDisableTiming);
    for (int k = 0; k < kAuxStreams && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&b->aux[k], hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&b->evjoin[k], hipEventDisableTiming);
    }
    return e;
}

// Launch fn(s, stream) for every system, system s on auxiliary stream
// s mod kAuxStreams, after the work already on b->stream and before what
// follows it (fork and join through events).
template <class F>
hipError_t per_system(utcnb_batch* b, F fn)
{
    hipError_t e = hipEventRecord(b->evfork, b->stream);
    for (int k = 0; k < kAuxStreams && e == hipSuccess; ++k)
        e = hipStreamWaitEvent(b->aux[k], b->evfork, 0);
    for (int s = 0; s < b->nsys && e == hipSuccess; ++s)
        e = fn(s, b->aux[s % kAuxStreams]);
    for (int k = 0; k < kAuxStreams && e == hipSuccess; ++k) {
        e = hipEventRecord(b->evjoin[k], b->aux[k]);
        if (e == hipSuccess)
            e = hipStreamWaitEvent(b->stream, b->evjoin[k], 0);
    }
    return e;
}

// Launch fn(g, stream) for every system group: group 0 on b->stream, group g
// on b->gstream[g], after the work already on b->stream and before what
// follows it (fork and join through events; in a captured graph the groups
// become independent branches).
template <class F>
hipError_t for_groups(utcnb_batch* b, F fn)
{
    if (b->ngroups == 1)
        return fn(0, b->stream);
    hipError_t e = hipEventRecord(b->gfork, b->stream);
    for (int g = 1; g < b->ngroups && e == hipSuccess; ++g)
        e = hipStreamWaitEvent(b->gstream[g], b->gfork, 0);
    for (int g = 0; g < b->ngroups && e == hipSuccess; ++g)
        e = fn(g, g == 0 ? b->stream : b->gstream[g]);
    for (int g = 1; g < b->ngroups && e == hipSuccess; ++g) {
        e = hipEventRecord(b->gjoin[g], b->gstream[g]);
        if (e == hipSuccess)
            e = hipStreamWaitEvent(b->stream, b->gjoin[g], 0);
    }
    return e;
}

hipError_t sym_step_group(const utcnb_batch* b, int sk, int g, const float4* so, float4* sn, double4* p,
                          double4* v, hipStream_t sg)
{
    return launch_batch_sym_step(b->sym_geom(sk, 0, g), b->snwaves[sk][g], b->group_bodies(g), so, sn, p, v,
                                 (float)b->cfg.eps2, b->cfg.dt, b->cfg.damping, b->cfg.kick, b->sipart,
                                 b->sjpart, sg);
}

// One run with the symmetric kernels (schedule sk: 0 persistent waves, 1 one
// wave per item): every system group runs its whole forward and adjoint on
// its own stream, touching only its own bodies and systems.
hipError_t enqueue_run_sym(utcnb_batch* b, int nsteps, int ckpt, bool adj, int sk)
{
    const double dt = b->cfg.dt, damp = b->cfg.damping;
    const int kick = b->cfg.kick;
    const size_t T = (size_t)b->total;
    const int k = ckpt < nsteps ? ckpt : nsteps;
    return for_groups(b, [&](int g, hipStream_t sg) -> hipError_t {
        const size_t b0 = (size_t)b->gbody[g], Tg = (size_t)b->group_bodies(g);
        const size_t sb = Tg * sizeof(double4);
        hipError_t e;
#define ENQ(x)                       \
    do {                             \
        if ((e = (x)) != hipSuccess) \
            return e;                \
    } while (0)
        // every run starts from the FP64 state, so a replayed graph is consistent
        ENQ(launch_snapshot(b->pos + b0, b->snap[0] + b0, (int)Tg, sg));
        int c = 0;
        const int nck = (nsteps + k - 1) / k, slast = (nck - 1) * k;
        for (int t = 0; t < nsteps; ++t) {
            if (adj && t % k == 0) {
                double4* ck = b->ckpt + 2 * T * (size_t)(t / k);
                ENQ(hipMemcpyAsync(ck + b0, b->pos + b0, sb, hipMemcpyDeviceToDevice, sg));
                ENQ(hipMemcpyAsync(ck + T + b0, b->vel + b0, sb, hipMemcpyDeviceToDevice, sg));
            }
            if (adj && t >= slast)  // the last segment's positions, kept instead of recomputed
                ENQ(hipMemcpyAsync(b->segpos + T * (size_t)(t - slast) + b0, b->pos + b0, sb,
                                   hipMemcpyDeviceToDevice, sg));
            ENQ(sym_step_group(b, sk, g, b->snap[c], b->snap[1 - c], b->pos, b->vel, sg));
            c = 1 - c;
        }
        if (!adj)
            return hipSuccess;
        const int s0g = b->gsys[g], nsg = b->gsys[g + 1] - s0g;
        const BatchGeom lg{b->d_off + s0g, b->d_n + s0g, b->d_chunks + s0g, nullptr, nullptr};
        ENQ(launch_batch_loss(lg, nsg, b->pos, b->tgt, b->xbar, b->vbar, b->mbar, b->loss + s0g, sg));
        for (int ci = nck - 1; ci >= 0; --ci) {
            const int s0 = ci * k, s1 = (s0 + k < nsteps) ? s0 + k : nsteps;
            const double4* ck = b->ckpt + 2 * T * (size_t)ci;
            if (ci != nck - 1) {  // recompute the segment's positions from its checkpoint
                ENQ(hipMemcpyAsync(b->wpos + b0, ck + b0, sb, hipMemcpyDeviceToDevice, sg));
                ENQ(hipMemcpyAsync(b->wvel + b0, ck + T + b0, sb, hipMemcpyDeviceToDevice, sg));
                ENQ(launch_snapshot(b->wpos + b0, b->wsnap[0] + b0, (int)Tg, sg));
                int wc = 0;
                for (int t = s0; t < s1; ++t) {
                    ENQ(hipMemcpyAsync(b->segpos + T * (size_t)(t - s0) + b0, b->wpos + b0, sb,
                                       hipMemcpyDeviceToDevice, sg));
                    if (t + 1 < s1) {
                        ENQ(sym_step_group(b, sk, g, b->wsnap[wc], b->wsnap[1 - wc], b->wpos, b->wvel, sg));
                        wc = 1 - wc;
                    }
                }
            }
            for (int t = s1 - 1; t >= s0; --t) {
                ENQ(launch_adjoint_pre(true, b->segpos + T * (size_t)(t - s0) + b0, nullptr, (int)Tg, dt, damp,
                                       kick, b->xbar + 3 * b0, b->vbar + 3 * b0, b->mbar + b0,
                                       b->jrec + 2 * b0, sg));
                ENQ(launch_batch_sym_vjp(b->sym_geom(sk, 1, g), b->snwaves[sk][g], (int)Tg, b->jrec,
                                         (float)b->cfg.eps2, b->svipart, b->svjpart, b->xbar, b->mbar, sg));
            }
        }
#undef ENQ
        return hipSuccess;
    });
}

// Enqueue one run (forward nsteps, optionally the adjoint) on b->stream.
// mode 0: batched kernels (one-pass with W = kWs[wi], or symmetric for
// wi >= kSymPersist); mode 1: per-system launches.
hipError_t enqueue_run(utcnb_batch* b, int nsteps, int ckpt, bool adj, int mode, int wi)
{
    hipStream_t st = b->stream;
    const float eps2 = (float)b->cfg.eps2;
    const double dt = b->cfg.dt, damp = b->cfg.damping;
    const int kick = b->cfg.kick;
    const size_t T = (size_t)b->total;
    const size_t sb = T * sizeof(double4);
    hipError_t e;
#define ENQ(x)                     \
    do {                           \
        if ((e = (x)) != hipSuccess) \
            return e;              \
    } while (0)

    if (mode == 0 && wi >= kSymPersist)
        return enqueue_run_sym(b, nsteps, ckpt, adj, wi - kSymPersist);
    auto step = [&](const float4* so, float4* sn, double4* p, double4* v) -> hipError_t {
        if (mode == 0)
            return launch_batch_step(kWs[wi], b->geom(wi), b->nblocks[wi], so, sn, p, v, eps2, dt,
                                     damp, kick, st);
        return per_system(b, [&](int s, hipStream_t ss) {
            StepArgs a{};
            a.n = b->n[s];
            a.n_pad = b->npad[s];
            a.eps2f = eps2;
            a.eps2 = b->cfg.eps2;
            a.dt = dt;
            a.damping = damp;
            a.kick = kick;
            a.newton = false;
            a.tile = per_system_tile(b->npad[s]);
            const size_t o = (size_t)b->off[s];
            return launch_mixed_step(a, so + o, sn + o, p + o, v + o, ss);
        });
    };
    auto vjp = [&]() -> hipError_t {
        if (mode == 0)
            return launch_batch_vjp(kWs[wi], b->geom(wi), b->nblocks[wi], b->jrec, eps2, b->xbar,
                                    b->mbar, st);
        return per_system(b, [&](int s, hipStream_t ss) {
            const size_t o = (size_t)b->off[s];
            return launch_vjp_f32(b->jrec + 2 * o, b->n[s], b->npad[s], b->cfg.eps2, b->xbar + 3 * o,
                                  b->mbar + o, b->cu_count, ss);
        });
    };

    // every run starts from the FP64 state, so a replayed graph is consistent
    ENQ(launch_snapshot(b->pos, b->snap[0], (int)T, st));
    int c = 0;
    const int k = ckpt < nsteps ? ckpt : nsteps;
    const int nck = (nsteps + k - 1) / k, slast = (nck - 1) * k;
    for (int t = 0; t < nsteps; ++t) {
        if (adj && t % k == 0) {
            double4* ck = b->ckpt + 2 * T * (size_t)(t / k);
            ENQ(hipMemcpyAsync(ck, b->pos, sb, hipMemcpyDeviceToDevice, st));
            ENQ(hipMemcpyAsync(ck + T, b->vel, sb, hipMemcpyDeviceToDevice, st));
        }
        if (adj && t >= slast)  // the last segment's positions, kept instead of recomputed
            ENQ(hipMemcpyAsync(b->segpos + T * (size_t)(t - slast), b->pos, sb, hipMemcpyDeviceToDevice, st));
        ENQ(step(b->snap[c], b->snap[1 - c], b->pos, b->vel));
        c = 1 - c;
    }
    if (!adj)
        return hipSuccess;
    ENQ(launch_batch_loss(b->geom(0), b->nsys, b->pos, b->tgt, b->xbar, b->vbar, b->mbar, b->loss, st));
    for (int ci = nck - 1; ci >= 0; --ci) {
        const int s0 = ci * k, s1 = (s0 + k < nsteps) ? s0 + k : nsteps;
        const double4* ck = b->ckpt + 2 * T * (size_t)ci;
        if (ci != nck - 1) {  // recompute the segment's positions from its checkpoint
            ENQ(hipMemcpyAsync(b->wpos, ck, sb, hipMemcpyDeviceToDevice, st));
            ENQ(hipMemcpyAsync(b->wvel, ck + T, sb, hipMemcpyDeviceToDevice, st));
            ENQ(launch_snapshot(b->wpos, b->wsnap[0], (int)T, st));
            int wc = 0;
            for (int t = s0; t < s1; ++t) {
                ENQ(hipMemcpyAsync(b->segpos + T * (size_t)(t - s0), b->wpos, sb,
                                   hipMemcpyDeviceToDevice, st));
                if (t + 1 < s1) {
                    ENQ(step(b->wsnap[wc], b->wsnap[1 - wc], b->wpos, b->wvel));
                    wc = 1 - wc;
                }
            }
        }
        for (int t = s1 - 1; t >= s0; --t) {
            ENQ(launch_adjoint_pre(true, b->segpos + T * (size_t)(t - s0), nullptr, (int)T, dt, damp,
                                   kick, b->xbar, b->vbar, b->mbar, b->jrec, st));
            ENQ(vjp());
        }
    }
#undef ENQ
    return hipSuccess;
}

}  // namespace

extern "C" {

int utcnb_batch_create(int nsys, const int32_t* sizes, const utcnb_config* cfg, void* stream,
                       utcnb_batch** out)
{
    if (!out || !sizes || !cfg || nsys < 1)
        return fail(UTCNB_EINVAL, "utcnb_batch_create: bad argument");
    *out = nullptr;
    if (cfg->precision != UTCNB_PREC_MIXED)
        return fail(UTCNB_EINVAL, "utcnb_batch_create: batched systems run in MIXED precision");
    if (cfg->kick != UTCNB_KICK_PHYSICAL && cfg->kick != UTCNB_KICK_REFERENCE)
        return fail(UTCNB_EINVAL, "utcnb_batch_create: unknown kick mode");
    if (!(cfg->eps2 > 0.0))
        return fail(UTCNB_EINVAL, "utcnb_batch_create: eps2 must be > 0");
    if (cfg->tile < 0 || cfg->tile > kSels)
        return fail(UTCNB_EINVAL, "utcnb_batch_create: tile must be 0 (tune) or 1..5");
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(UTCNB_ENODEV, "utcnb_batch_create: no HIP device");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(UTCNB_ENODEV, "utcnb_batch_create: built for gfx950 only");
    utcnb_batch* b = new (std::nothrow) utcnb_batch();
    if (!b)
        return fail(UTCNB_ENOMEM, "utcnb_batch_create: out of host memory");
    b->cfg = *cfg;
    b->cu_count = prop.multiProcessorCount;
    b->partial_budget = sym_partial_budget(prop.totalGlobalMem);
    b->nsys = nsys;
    // the symmetric kernels (candidates when tuning) need systems in whole wave blocks
    b->align = (cfg->tile == 0 || cfg->tile > kSymPersist) ? kBatchSymAlign : kChunk;
    long long o = 0;
    for (int s = 0; s < nsys; ++s) {
        if (sizes[s] < 1 || sizes[s] > (1 << 24)) {
            delete b;
            return fail(UTCNB_EINVAL, "utcnb_batch_create: system sizes must be in [1, 2^24]");
        }
        b->n.push_back(sizes[s]);
        b->npad.push_back((sizes[s] + b->align - 1) / b->align * b->align);
        b->off.push_back((int)o);
        o += b->npad.back();
        b->total_real += sizes[s];
    }
    if (o > (1LL << 30)) {
        delete b;
        return fail(UTCNB_EINVAL, "utcnb_batch_create: too many bodies in total");
    }
    b->total = o;
    int rc = UTCNB_OK;
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess) {
        b->own_stream = true;  // capture needs a stream other than the null stream
    } else {
        rc = fail(UTCNB_EHIP, "utcnb_batch_create: hipStreamCreate failed");
    }
    const size_t T = (size_t)b->total;
    if (!rc)
        rc = dalloc(&b->pos, T) || dalloc(&b->vel, T) || dalloc(&b->snap[0], T) ||
             dalloc(&b->snap[1], T) || dalloc(&b->d_off, (size_t)nsys) ||
             dalloc(&b->d_n, (size_t)nsys) || dalloc(&b->d_chunks, (size_t)nsys);
    if (!rc) {
        std::vector<int> chunks(nsys);
        for (int s = 0; s < nsys; ++s)
            chunks[s] = b->npad[s] / kChunk;
        hipError_t e = hipMemcpy(b->d_off, b->off.data(), sizeof(int) * nsys, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_n, b->n.data(), sizeof(int) * nsys, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_chunks, chunks.data(), sizeof(int) * nsys, hipMemcpyHostToDevice);
        for (int wi = 0; wi < 3 && e == hipSuccess; ++wi) {
            long long nitems = 0;
            for (int s = 0; s < nsys; ++s)
                nitems += (b->n[s] + 63) / 64;
            const int W = kWs[wi];
            int blocks = (int)std::min<long long>(nitems, (long long)b->cu_count * 32 / W);
            std::vector<int2> items;
            std::vector<int> start;
            schedule(b, W, blocks, items, start);
            b->nblocks[wi] = blocks;
            if (dalloc(&b->d_items[wi], items.size()) || dalloc(&b->d_start[wi], start.size())) {
                e = hipErrorOutOfMemory;
                break;
            }
            e = hipMemcpy(b->d_items[wi], items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(b->d_start[wi], start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess && b->align == kBatchSymAlign)
            e = setup_sym(b);
        if (e == hipSuccess && cfg->tile > kSymPersist && !b->sym_ok)
            rc = fail(UTCNB_EINVAL, "utcnb_batch_create: the symmetric batched kernels' partial sums "
                                    "would exceed 16 GiB for these systems; use tile 0..3");
        if (e == hipSuccess)
            e = launch_batch_pad(b->geom(0), nsys, b->pos, b->vel, b->snap[0], b->snap[1], b->stream);
        if (e == hipSuccess && (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess))
            e = hipErrorUnknown;
        if (e == hipSuccess)
            e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess && !rc)
            rc = fail(UTCNB_EHIP, std::string("utcnb_batch_create: ") + hipGetErrorString(e));
    }
    if (cfg->tile > 0)
        b->wsel = cfg->tile - 1;
    if (rc) {
        release(b);
        delete b;
        return rc < 0 ? rc : UTCNB_ENOMEM;
    }
    *out = b;
    return UTCNB_OK;
}

int utcnb_batch_destroy(utcnb_batch* b)
{
    if (!b)
        return UTCNB_OK;
    (void)hipStreamSynchronize(b->stream);
    release(b);
    delete b;
    return UTCNB_OK;
}

int utcnb_batch_set_state(utcnb_batch* b, const double* pos4, const double* vel4)
{
    if (!b || !pos4 || !vel4)
        return fail(UTCNB_EINVAL, "utcnb_batch_set_state: null argument");
    const size_t T = (size_t)b->total;
    std::vector<double> hp(4 * T), hv(4 * T);
    size_t src = 0;
    for (int s = 0; s < b->nsys; ++s) {
        const size_t o = (size_t)b->off[s];
        for (int t = 0; t < b->npad[s]; ++t) {
            double* p = &hp[4 * (o + t)];
            double* v = &hv[4 * (o + t)];
            if (t < b->n[s]) {
                std::memcpy(p, pos4 + 4 * src, 4 * sizeof(double));
                std::memcpy(v, vel4 + 4 * src, 4 * sizeof(double));
                ++src;
            } else {
                p[0] = p[1] = p[2] = kPadCoord;
                p[3] = 0.0;
                v[0] = v[1] = v[2] = 0.0;
                v[3] = 1.0;
            }
        }
    }
    UTCNB_HIP(hipMemcpyAsync(b->pos, hp.data(), 4 * T * sizeof(double), hipMemcpyHostToDevice, b->stream));
    UTCNB_HIP(hipMemcpyAsync(b->vel, hv.data(), 4 * T * sizeof(double), hipMemcpyHostToDevice, b->stream));
    UTCNB_HIP(hipStreamSynchronize(b->stream));
    b->has_state = true;
    return UTCNB_OK;
}

int utcnb_batch_get_state(utcnb_batch* b, double* pos4, double* vel4)
{
    if (!b)
        return fail(UTCNB_EINVAL, "utcnb_batch_get_state: null handle");
    if (!b->has_state)
        return fail(UTCNB_ESTATE, "utcnb_batch_get_state: no state set");
    const size_t T = (size_t)b->total;
    std::vector<double> hp(4 * T), hv(4 * T);
    UTCNB_HIP(hipMemcpyAsync(hp.data(), b->pos, 4 * T * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    UTCNB_HIP(hipMemcpyAsync(hv.data(), b->vel, 4 * T * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    UTCNB_HIP(hipStreamSynchronize(b->stream));
    size_t dst = 0;
    for (int s = 0; s < b->nsys; ++s)
        for (int t = 0; t < b->n[s]; ++t, ++dst) {
            const size_t i = (size_t)b->off[s] + t;
            if (pos4)
                std::memcpy(pos4 + 4 * dst, &hp[4 * i], 4 * sizeof(double));
            if (vel4)
                std::memcpy(vel4 + 4 * dst, &hv[4 * i], 4 * sizeof(double));
        }
    return UTCNB_OK;
}

// Pick the batched kernel (one-pass W = 4, 8, 16; symmetric persistent or one
// wave per item when the layout allows) by timing forward steps of each on a
// copy of the state.
static int tune_w(utcnb_batch* b)
{
    const size_t T = (size_t)b->total;
    double4 *sp = nullptr, *sv = nullptr;
    if (dalloc(&sp, T) || dalloc(&sv, T)) {
        dfree(sp);
        return fail(UTCNB_ENOMEM, "utcnb_batch_run: tuning buffers");
    }
    UTCNB_HIP(hipMemcpyAsync(sp, b->pos, T * sizeof(double4), hipMemcpyDeviceToDevice, b->stream));
    UTCNB_HIP(hipMemcpyAsync(sv, b->vel, T * sizeof(double4), hipMemcpyDeviceToDevice, b->stream));
    float best = 1e30f;
    int bw = 1;
    const float eps2 = (float)b->cfg.eps2;
    auto one = [&](int wi) -> hipError_t {
        if (wi >= kSymPersist)
            return for_groups(b, [&](int g, hipStream_t sg) {
                return sym_step_group(b, wi - kSymPersist, g, b->snap[0], b->snap[1], b->pos, b->vel, sg);
            });
        return launch_batch_step(kWs[wi], b->geom(wi), b->nblocks[wi], b->snap[0], b->snap[1], b->pos,
                                 b->vel, eps2, b->cfg.dt, b->cfg.damping, b->cfg.kick, b->stream);
    };
    for (int wi = 0; wi < (b->has_sym() ? kSels : 3); ++wi) {
        UTCNB_HIP(one(wi));
        UTCNB_HIP(hipEventRecord(b->ev0, b->stream));
        for (int r = 0; r < 3; ++r)
            UTCNB_HIP(one(wi));
        UTCNB_HIP(hipEventRecord(b->ev1, b->stream));
        UTCNB_HIP(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        UTCNB_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        if (ms < best) {
            best = ms;
            bw = wi;
        }
    }
    UTCNB_HIP(hipMemcpyAsync(b->pos, sp, T * sizeof(double4), hipMemcpyDeviceToDevice, b->stream));
    UTCNB_HIP(hipMemcpyAsync(b->vel, sv, T * sizeof(double4), hipMemcpyDeviceToDevice, b->stream));
    UTCNB_HIP(hipStreamSynchronize(b->stream));
    dfree(sp);
    dfree(sv);
    b->wsel = bw;
    return UTCNB_OK;
}

int utcnb_batch_get_tile(const utcnb_batch* b, int* tile)
{
    if (!b || !tile)
        return fail(UTCNB_EINVAL, "utcnb_batch_get_tile: null argument");
    *tile = b->wsel + 1;
    return UTCNB_OK;
}

int utcnb_batch_run(utcnb_batch* b, int nsteps, int mode, int ckpt_every, const double* target3,
                    double* losses, double* dx0, double* dv0, double* dm, double* ms)
{
    if (!b || nsteps < 1 || mode < 0 || mode > 2)
        return fail(UTCNB_EINVAL, "utcnb_batch_run: bad argument");
    if (!b->has_state)
        return fail(UTCNB_ESTATE, "utcnb_batch_run: no state set");
    const bool adj = target3 != nullptr;
    if (adj) {
        if (!losses || !dx0 || !dv0 || !dm || ckpt_every < 1)
            return fail(UTCNB_EINVAL, "utcnb_batch_run: adjoint outputs and ckpt_every >= 1 required");
        if (b->cfg.kick != UTCNB_KICK_PHYSICAL)
            return fail(UTCNB_EINVAL, "utcnb_batch_run: the batched adjoint supports the physical kick");
        const int k = ckpt_every < nsteps ? ckpt_every : nsteps;
        int rc = adj_workspace(b, (nsteps + k - 1) / k, k);
        if (rc)
            return rc;
        // targets, compact host layout -> padded device layout
        const size_t T = (size_t)b->total;
        std::vector<double> ht(3 * T, 0.0);
        size_t src = 0;
        for (int s = 0; s < b->nsys; ++s)
            for (int t = 0; t < b->n[s]; ++t, ++src)
                std::memcpy(&ht[3 * ((size_t)b->off[s] + t)], target3 + 3 * src, 3 * sizeof(double));
        UTCNB_HIP(hipMemcpyAsync(b->tgt, ht.data(), 3 * T * sizeof(double), hipMemcpyHostToDevice, b->stream));
        UTCNB_HIP(hipStreamSynchronize(b->stream));
    }
    if (mode != 1 && b->wsel < 0) {
        int rc = tune_w(b);
        if (rc)
            return rc;
    }
    const int wi = b->wsel < 0 ? 1 : b->wsel;
    if (mode == 0) {
        if (!b->exec || b->g_nsteps != nsteps || b->g_ckpt != ckpt_every || b->g_adj != adj ||
            b->g_w != wi) {
            if (b->exec)
                (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
            hipGraph_t graph = nullptr;
            UTCNB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_run(b, nsteps, adj ? ckpt_every : nsteps, adj, 0, wi);
            hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (graph)
                    (void)hipGraphDestroy(graph);
                return fail(UTCNB_EHIP, std::string("utcnb_batch_run: capture: ") +
                                            hipGetErrorString(e != hipSuccess ? e : e2));
            }
            e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess)
                return fail(UTCNB_EHIP, std::string("utcnb_batch_run: instantiate: ") + hipGetErrorString(e));
            b->g_nsteps = nsteps;
            b->g_ckpt = ckpt_every;
            b->g_adj = adj;
            b->g_w = wi;
        }
        UTCNB_HIP(hipEventRecord(b->ev0, b->stream));
        UTCNB_HIP(hipGraphLaunch(b->exec, b->stream));
        UTCNB_HIP(hipEventRecord(b->ev1, b->stream));
    } else {
        // mode 1: per-system launches; mode 2: the persistent kernels, eagerly
        if (mode == 1)
            UTCNB_HIP(aux_streams(b));
        UTCNB_HIP(hipEventRecord(b->ev0, b->stream));
        hipError_t e = enqueue_run(b, nsteps, adj ? ckpt_every : nsteps, adj, mode == 2 ? 0 : 1, wi);
        if (e != hipSuccess)
            return fail(UTCNB_EHIP, std::string("utcnb_batch_run: ") + hipGetErrorString(e));
        UTCNB_HIP(hipEventRecord(b->ev1, b->stream));
    }
    UTCNB_HIP(hipEventSynchronize(b->ev1));
    if (ms) {
        float f = 0.f;
        UTCNB_HIP(hipEventElapsedTime(&f, b->ev0, b->ev1));
        *ms = f;
    }
    if (adj) {
        const size_t T = (size_t)b->total;
        std::vector<double> hx(3 * T), hv(3 * T), hm(T);
        UTCNB_HIP(hipMemcpyAsync(hx.data(), b->xbar, 3 * T * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        UTCNB_HIP(hipMemcpyAsync(hv.data(), b->vbar, 3 * T * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        UTCNB_HIP(hipMemcpyAsync(hm.data(), b->mbar, T * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        UTCNB_HIP(hipMemcpyAsync(losses, b->loss, b->nsys * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        UTCNB_HIP(hipStreamSynchronize(b->stream));
        size_t dst = 0;
        for (int s = 0; s < b->nsys; ++s)
            for (int t = 0; t < b->n[s]; ++t, ++dst) {
                const size_t i = (size_t)b->off[s] + t;
                std::memcpy(dx0 + 3 * dst, &hx[3 * i], 3 * sizeof(double));
                std::memcpy(dv0 + 3 * dst, &hv[3 * i], 3 * sizeof(double));
                dm[dst] = hm[i];
            }
    }
    return UTCNB_OK;
}

}  // extern "C"
