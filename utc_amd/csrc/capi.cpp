// capi.cpp -- the C ABI (include/utcnb.h): system handles over the gfx950 kernels.
//
// A handle owns the device state of one system in the layout its precision
// needs (citations relative to the reference's utc-0.1/ tree):
//   REFERENCE f32/f64 : pos T4[2] (double buffered) + vel T4, like the
//                       GpuData pair of benchapps/nbody/utc/sgpu/body_task_sgpu.cu:100-101
//   FP32              : pos float4[2] padded to kPadJ + vel float4
//   MIXED             : FP64 pos/vel double4 (updated in place) + FP32
//                       snapshot float4[2] padded, read by the force kernel
//   FP64              : pos double4[2] padded + vel double4
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

namespace utcnb {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace utcnb

using namespace utcnb;

struct utcnb_system {
    utcnb_config cfg{};
    hipStream_t stream = nullptr;
    int device = 0;
    int n = 0, n_pad = 0;
    int cur = 0;
    bool has_state = false;
    StepArgs args{};
    float4* p32[2] = {nullptr, nullptr};
    float4* v32 = nullptr;
    double4* p64[2] = {nullptr, nullptr};
    double4* v64 = nullptr;
    float4* snap[2] = {nullptr, nullptr};
    double* scratch = nullptr;   // 3n doubles (field) / energy partials
    float* ipart = nullptr;      // symmetric-tile workspace
    float* jpart = nullptr;
    float* share_ipart = nullptr;  // target-side partials of a share step (utcnb_share_pairs)
    size_t share_ipart_floats = 0;
    float* cipart = nullptr;     // chunked symmetric steps: target-side partials,
    float* cjpart = nullptr;     //   one chunk's reaction partials,
    double* react = nullptr;     //   reactions summed per body
    size_t scratch_bytes = 0;
    void* stage = nullptr;       // conversion staging, 4n * 8 bytes
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int last_launches = 0;
    // adjoint workspace (allocated on first use, grown on demand)
    struct Adj {
        int nck = 0, seg = 0;
        double4* ckpt = nullptr;     // nck * 2n (pos, vel)
        double4* segpos = nullptr;   // seg * n
        double4* segpos2 = nullptr;  // seg * n, second segment buffer (nck > 1)
        hipStream_t rstream = nullptr;  // segment recomputes beside the VJP steps
        hipEvent_t evfwd = nullptr, evrec[2] = {nullptr, nullptr}, evvjp = nullptr;
        double4* wpos[2] = {nullptr, nullptr};  // n_pad (FP64 work positions)
        double4* wvel = nullptr;     // n
        float4* wsnap[2] = {nullptr, nullptr};  // n_pad (MIXED work snapshots)
        void* jrec = nullptr;        // 2 * n_pad records (float4 or double4)
        void* fsrc = nullptr;        // n_pad padded field source (float4 or double4)
        double* xbar = nullptr;      // 3n
        double* vbar = nullptr;      // 3n
        double* mbar = nullptr;      // n
        double* tgt = nullptr;       // 3n
        double* part = nullptr;      // loss partials + loss
        int vjp_r = 0;               // symmetric VJP: waves per wave block (0 = one-pass VJP)
        float4* vjp_ipart = nullptr; // symmetric VJP partials
        float4* vjp_jpart = nullptr;
    } adj;
    bool timed = false;
    int cu_count = 256;
    size_t partial_budget = (size_t)16 << 30;  // sym_partial_budget of the device
};

namespace {

int check_device(int* dev, int* cus, size_t* mem = nullptr)
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess)
        return fail(UTCNB_ENODEV, "no HIP device available");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) != hipSuccess)
        return fail(UTCNB_ENODEV, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(UTCNB_ENODEV, std::string("device is ") + prop.gcnArchName +
                                      ", this library is built for gfx950 only");
    *dev = d;
    if (cus)
        *cus = prop.multiProcessorCount;
    if (mem)
        *mem = prop.totalGlobalMem;
    return UTCNB_OK;
}

template <typename T>
int dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0)
        return UTCNB_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(UTCNB_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
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

void release_adjoint(utcnb_system* s)
{
    auto& a = s->adj;
    dfree(a.ckpt);
    dfree(a.segpos);
    dfree(a.segpos2);
    for (hipEvent_t* ev : {&a.evfwd, &a.evrec[0], &a.evrec[1], &a.evvjp}) {
        if (*ev)
            (void)hipEventDestroy(*ev);
        *ev = nullptr;
    }
    if (a.rstream)
        (void)hipStreamDestroy(a.rstream);
    a.rstream = nullptr;
    dfree(a.wpos[0]);
    dfree(a.wpos[1]);
    dfree(a.wvel);
    dfree(a.wsnap[0]);
    dfree(a.wsnap[1]);
    if (a.jrec)
        (void)hipFree(a.jrec);
    if (a.fsrc)
        (void)hipFree(a.fsrc);
    a.jrec = a.fsrc = nullptr;
    dfree(a.xbar);
    dfree(a.vbar);
    dfree(a.mbar);
    dfree(a.tgt);
    dfree(a.part);
    dfree(a.vjp_ipart);
    dfree(a.vjp_jpart);
    a.nck = a.seg = a.vjp_r = 0;
}

void release(utcnb_system* s)
{
    release_adjoint(s);
    dfree(s->p32[0]);
    dfree(s->p32[1]);
    dfree(s->v32);
    dfree(s->p64[0]);
    dfree(s->p64[1]);
    dfree(s->v64);
    dfree(s->snap[0]);
    dfree(s->snap[1]);
    dfree(s->scratch);
    dfree(s->ipart);
    dfree(s->jpart);
    dfree(s->share_ipart);
    s->share_ipart_floats = 0;
    dfree(s->cipart);
    dfree(s->cjpart);
    dfree(s->react);
    if (s->stage)
        (void)hipFree(s->stage);
    s->stage = nullptr;
    if (s->ev0)
        (void)hipEventDestroy(s->ev0);
    if (s->ev1)
        (void)hipEventDestroy(s->ev1);
}

bool is_f32_state(const utcnb_system* s)
{
    return s->cfg.precision == UTCNB_PREC_FP32 ||
           (s->cfg.precision == UTCNB_PREC_REFERENCE && s->cfg.state_dtype == UTCNB_F32);
}

// Current FP64 position / velocity arrays (null when the state is FP32).
double4* cur_p64(utcnb_system* s)
{
    return s->cfg.precision == UTCNB_PREC_MIXED ? s->p64[0] : s->p64[s->cur];
}

hipMemcpyKind kind_of(int where, bool to_device)
{
    if (where == UTCNB_DEVICE)
        return hipMemcpyDeviceToDevice;
    return to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
}

// Bring an input array (n*4 of dtype, host or device) to the device as `T4`.
template <typename T4>
int upload4(utcnb_system* s, const void* src, int dtype, int where, T4* dst)
{
    const size_t n4 = 4 * (size_t)s->n;
    constexpr bool dst_f32 = sizeof(T4) == sizeof(float4);
    const bool same = (dtype == UTCNB_F32) == dst_f32;
    if (same) {
        UTCNB_HIP(hipMemcpyAsync(dst, src, n4 * (dst_f32 ? 4 : 8), kind_of(where, true),
                                 s->stream));
        return UTCNB_OK;
    }
    const size_t sb = n4 * (dtype == UTCNB_F32 ? 4 : 8);
    UTCNB_HIP(hipMemcpyAsync(s->stage, src, sb, kind_of(where, true), s->stream));
    hipError_t e;
    if constexpr (dst_f32)
        e = launch_f64_to_f32((const double4*)s->stage, (float4*)dst, s->n, s->stream);
    else
        e = launch_f32_to_f64((const float4*)s->stage, (double4*)dst, s->n, s->stream);
    UTCNB_HIP(e);
    return UTCNB_OK;
}

template <typename T4>
int download4(utcnb_system* s, const T4* src, void* dst, int dtype, int where)
{
    const size_t n4 = 4 * (size_t)s->n;
    constexpr bool src_f32 = sizeof(T4) == sizeof(float4);
    const bool same = (dtype == UTCNB_F32) == src_f32;
    if (same) {
        UTCNB_HIP(hipMemcpyAsync(dst, src, n4 * (src_f32 ? 4 : 8), kind_of(where, false),
                                 s->stream));
        return UTCNB_OK;
    }
    hipError_t e;
    if constexpr (src_f32)
        e = launch_f32_to_f64((const float4*)src, (double4*)s->stage, s->n, s->stream);
    else
        e = launch_f64_to_f32((const double4*)src, (float4*)s->stage, s->n, s->stream);
    UTCNB_HIP(e);
    const size_t db = n4 * (dtype == UTCNB_F32 ? 4 : 8);
    UTCNB_HIP(hipMemcpyAsync(dst, s->stage, db, kind_of(where, false), s->stream));
    return UTCNB_OK;
}

}  // namespace

extern "C" {

const char* utcnb_last_error(void) { return g_last_error.c_str(); }

const char* utcnb_version(void) { return "utcnb 0.1 gfx950"; }

int utcnb_device_info(char* name, int name_len, int* cu_count, int* clock_khz)
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess)
        return fail(UTCNB_ENODEV, "no HIP device available");
    hipDeviceProp_t prop;
    UTCNB_HIP(hipGetDeviceProperties(&prop, d));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", prop.gcnArchName);
    }
    if (cu_count)
        *cu_count = prop.multiProcessorCount;
    if (clock_khz)
        *clock_khz = prop.clockRate;
    return UTCNB_OK;
}

int utcnb_integrate_f32(const float* pos_old, float* pos_new, float* vel, int numBodies,
                        float deltaTime, float softeningSquared, float damping, int blocksize,
                        void* stream)
{
    if (!pos_old || !pos_new || !vel || numBodies <= 0)
        return fail(UTCNB_EINVAL, "utcnb_integrate_f32: null pointer or numBodies <= 0");
    if (pos_old == pos_new)
        return fail(UTCNB_EINVAL, "utcnb_integrate_f32: pos_old and pos_new alias");
    if (blocksize != 0 && (blocksize < 64 || blocksize > 1024 || blocksize % 64 != 0))
        return fail(UTCNB_EINVAL, "utcnb_integrate_f32: blocksize must be 64..1024, multiple of 64");
    UTCNB_HIP(launch_ref_step_f32((const float4*)pos_old, (float4*)pos_new, (float4*)vel,
                                  numBodies, deltaTime, softeningSquared, damping, blocksize,
                                  (hipStream_t)stream));
    return UTCNB_OK;
}

int utcnb_integrate_f64(const double* pos_old, double* pos_new, double* vel, int numBodies,
                        double deltaTime, double softeningSquared, double damping,
                        int blocksize, void* stream)
{
    if (!pos_old || !pos_new || !vel || numBodies <= 0)
        return fail(UTCNB_EINVAL, "utcnb_integrate_f64: null pointer or numBodies <= 0");
    if (pos_old == pos_new)
        return fail(UTCNB_EINVAL, "utcnb_integrate_f64: pos_old and pos_new alias");
    if (blocksize != 0 && (blocksize < 64 || blocksize > 1024 || blocksize % 64 != 0))
        return fail(UTCNB_EINVAL, "utcnb_integrate_f64: blocksize must be 64..1024, multiple of 64");
    UTCNB_HIP(launch_ref_step_f64((const double4*)pos_old, (double4*)pos_new, (double4*)vel,
                                  numBodies, deltaTime, softeningSquared, damping, blocksize,
                                  (hipStream_t)stream));
    return UTCNB_OK;
}

static int check_slice_args(const char* who, int numBodies, int start, int count)
{
    if (numBodies <= 0 || start < 0 || count < 1 || (int64_t)start + count > numBodies)
        return fail(UTCNB_EINVAL, std::string(who) + ": need 0 <= start, count >= 1, start + count <= numBodies");
    return UTCNB_OK;
}

int utcnb_integrate_slice_f32(const float* pos_old, float* pos_new, float* vel, int numBodies, int start,
                              int count, float deltaTime, float softeningSquared, float damping, int blocksize,
                              void* stream)
{
    if (!pos_old || !pos_new || !vel)
        return fail(UTCNB_EINVAL, "utcnb_integrate_slice_f32: null pointer");
    if (int rc = check_slice_args("utcnb_integrate_slice_f32", numBodies, start, count))
        return rc;
    if (blocksize != 0 && (blocksize < 64 || blocksize > 1024 || blocksize % 64 != 0))
        return fail(UTCNB_EINVAL, "utcnb_integrate_slice_f32: blocksize must be 64..1024, multiple of 64");
    UTCNB_HIP(launch_ref_step_slice_f32((const float4*)pos_old, (float4*)pos_new, (float4*)vel, numBodies, start,
                                        count, deltaTime, softeningSquared, damping, blocksize,
                                        (hipStream_t)stream));
    return UTCNB_OK;
}

int utcnb_integrate_slice_f64(const double* pos_old, double* pos_new, double* vel, int numBodies, int start,
                              int count, double deltaTime, double softeningSquared, double damping,
                              int blocksize, void* stream)
{
    if (!pos_old || !pos_new || !vel)
        return fail(UTCNB_EINVAL, "utcnb_integrate_slice_f64: null pointer");
    if (int rc = check_slice_args("utcnb_integrate_slice_f64", numBodies, start, count))
        return rc;
    if (blocksize != 0 && (blocksize < 64 || blocksize > 1024 || blocksize % 64 != 0))
        return fail(UTCNB_EINVAL, "utcnb_integrate_slice_f64: blocksize must be 64..1024, multiple of 64");
    UTCNB_HIP(launch_ref_step_slice_f64((const double4*)pos_old, (double4*)pos_new, (double4*)vel, numBodies,
                                        start, count, deltaTime, softeningSquared, damping, blocksize,
                                        (hipStream_t)stream));
    return UTCNB_OK;
}

int utcnb_tile_count(void) { return tile_count(); }

const char* utcnb_tile_name(int tile)
{
    if (tile < 1 || tile > tile_count())
        return "";
    return tile_get(tile - 1).name;
}

int utcnb_create(const utcnb_config* cfg, void* stream, utcnb_system** out)
{
    if (!cfg || !out)
        return fail(UTCNB_EINVAL, "utcnb_create: null argument");
    *out = nullptr;
    const utcnb_config c = *cfg;
    if (c.n < 1 || (int64_t)c.n > (int64_t)1 << 30)
        return fail(UTCNB_EINVAL, "utcnb_create: n must be in [1, 2^30]");
    if (c.precision < UTCNB_PREC_REFERENCE || c.precision > UTCNB_PREC_FP64)
        return fail(UTCNB_EINVAL, "utcnb_create: unknown precision");
    if (c.kick != UTCNB_KICK_REFERENCE && c.kick != UTCNB_KICK_PHYSICAL)
        return fail(UTCNB_EINVAL, "utcnb_create: unknown kick mode");
    if (c.precision == UTCNB_PREC_REFERENCE) {
        if (c.state_dtype != UTCNB_F32 && c.state_dtype != UTCNB_F64)
            return fail(UTCNB_EINVAL, "utcnb_create: state_dtype must be F32 or F64");
        if (c.kick != UTCNB_KICK_REFERENCE)
            return fail(UTCNB_EINVAL, "utcnb_create: REFERENCE precision uses the reference kick");
        if (!(c.eps2 >= 0.0))
            return fail(UTCNB_EINVAL, "utcnb_create: eps2 must be >= 0");
    } else if (!(c.eps2 > 0.0) || !std::isfinite(c.eps2)) {
        return fail(UTCNB_EINVAL, "utcnb_create: eps2 must be finite and > 0");
    }
    if (!std::isfinite(c.dt) || !std::isfinite(c.damping))
        return fail(UTCNB_EINVAL, "utcnb_create: dt and damping must be finite");
    if (c.tile < 0 || c.tile > tile_count())
        return fail(UTCNB_EINVAL, "utcnb_create: tile out of range");

    int dev = 0, cus = 256;
    size_t mem = 0;
    int rc = check_device(&dev, &cus, &mem);
    if (rc)
        return rc;

    utcnb_system* s = new (std::nothrow) utcnb_system();
    if (!s)
        return fail(UTCNB_ENOMEM, "utcnb_create: out of host memory");
    s->cfg = c;
    s->stream = (hipStream_t)stream;
    s->device = dev;
    s->cu_count = cus;
    s->partial_budget = sym_partial_budget(mem);
    s->n = c.n;
    const bool fast = c.precision != UTCNB_PREC_REFERENCE;
    s->n_pad = fast ? (int)pad_j(c.n) : c.n;
    const size_t np = (size_t)s->n_pad, n = (size_t)c.n;

    s->args.n = c.n;
    s->args.n_pad = s->n_pad;
    s->args.eps2 = c.eps2;
    s->args.eps2f = (float)c.eps2;
    s->args.dt = c.dt;
    s->args.damping = c.damping;
    s->args.kick = c.kick;
    s->args.newton = (c.flags & UTCNB_FLAG_NO_NEWTON) == 0;
    s->args.tile = c.tile > 0 ? c.tile - 1 : tile_default(c.n, cus, s->partial_budget);
    if (c.tile == 0 && c.precision == UTCNB_PREC_MIXED && !tile_get(s->args.tile).sym) {
        // a system whose symmetric partial sums exceed the budget in one
        // launch still takes the symmetric tile, evaluated in chunks
        const int t = tile_default(c.n, cus, SIZE_MAX);
        if (tile_get(t).sym)
            s->args.tile = t;
    }
    if (fast && c.precision != UTCNB_PREC_FP64 && tile_get(s->args.tile).cap > 0 &&
        s->n_pad > tile_get(s->args.tile).cap) {
        delete s;
        return fail(UTCNB_EINVAL, std::string("utcnb_create: tile ") + tile_get(s->args.tile).name + " holds at most " +
                                      std::to_string(tile_get(s->args.tile).cap) + " bodies (N padded to 1024)");
    }

    switch (c.precision) {
    case UTCNB_PREC_REFERENCE:
        if (c.state_dtype == UTCNB_F32)
            rc = dalloc(&s->p32[0], np) || dalloc(&s->p32[1], np) || dalloc(&s->v32, n);
        else
            rc = dalloc(&s->p64[0], np) || dalloc(&s->p64[1], np) || dalloc(&s->v64, n);
        break;
    case UTCNB_PREC_FP32:
        rc = dalloc(&s->p32[0], np) || dalloc(&s->p32[1], np) || dalloc(&s->v32, n);
        break;
    case UTCNB_PREC_MIXED:
        rc = dalloc(&s->p64[0], n) || dalloc(&s->v64, n) || dalloc(&s->snap[0], np) ||
             dalloc(&s->snap[1], np);
        break;
    case UTCNB_PREC_FP64:
        rc = dalloc(&s->p64[0], np) || dalloc(&s->p64[1], np) || dalloc(&s->v64, n);
        break;
    }
    if (!rc && (c.precision == UTCNB_PREC_FP32 || c.precision == UTCNB_PREC_MIXED) &&
        tile_get(s->args.tile).sym) {
        // symmetric tiles keep one reaction partial per (wave block, wave block) pair
        const Tile& t = tile_get(s->args.tile);
        if (np % (size_t)(128 * t.k) != 0) {
            rc = fail(UTCNB_EINVAL, "utcnb_create: symmetric tile needs N padded to 128*K");
        } else {
            const size_t jf = sym_jpart_floats((int)np, t.k, t.rs),
                         jif = sym_ipart_floats((int)np, t.k, sym_tile_r(t, (int)np));
            const bool full_fits = (jf + jif) * sizeof(float) <= s->partial_budget;
            if (c.precision == UTCNB_PREC_MIXED && (!full_fits || (c.flags & UTCNB_FLAG_SYM_CHUNKS))) {
                // chunks of wave blocks: the fewest (at least 2) whose partial
                // sums and per-body reaction sums fit the budget
                const int NW = (int)(np / (64 * (size_t)t.k));
                int G = 2;
                size_t cif = 0, cjf = 0;
                int Rs = 0;
                for (; G <= NW; ++G) {
                    const int nw = (NW + G - 1) / G;
                    Rs = sym_share_splits(nw, t.s, t.k, NW);
                    cif = (size_t)Rs * np * 3;
                    cjf = sym_chunk_jpart_floats(nw, NW, t.k);
                    if ((cif + cjf) * sizeof(float) + 3 * n * sizeof(double) <= s->partial_budget)
                        break;
                }
                if (G > NW) {
                    rc = fail(UTCNB_EINVAL, "utcnb_create: N too large for a symmetric tile even in chunks");
                } else {
                    rc = dalloc(&s->cipart, cif) || dalloc(&s->cjpart, cjf) || dalloc(&s->react, 3 * n);
                    s->args.chunks = G;
                    s->args.chunk_r = Rs;
                    s->args.cipart = s->cipart;
                    s->args.cjpart = s->cjpart;
                    s->args.react = s->react;
                }
            } else if (!full_fits) {
                rc = fail(UTCNB_EINVAL, "utcnb_create: N too large for a symmetric tile (partial sums "
                                        "above a quarter of the device, at most 64 GiB); use a one-pass tile "
                                        "(or MIXED, which evaluates the symmetric step in chunks)");
            } else {
                rc = dalloc(&s->ipart, jif) || dalloc(&s->jpart, jf);
            }
        }
        s->args.ipart = s->ipart;
        s->args.jpart = s->jpart;
    }
    if (!rc) {
        s->scratch_bytes = sizeof(double) * (3 * n + 5 + 5 * ((n + 255) / 256));
        rc = dalloc((char**)&s->scratch, s->scratch_bytes);
    }
    if (!rc) {
        hipError_t e = hipMalloc(&s->stage, 4 * n * sizeof(double) + 4 * n * sizeof(double));
        if (e != hipSuccess)
            rc = fail(UTCNB_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    if (!rc && (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess))
        rc = fail(UTCNB_EHIP, "hipEventCreate failed");
    if (!rc && fast) {
        hipError_t e = hipSuccess;
        for (int b = 0; b < 2 && e == hipSuccess; ++b) {
            if (s->p32[b])
                e = launch_pad_f32(s->p32[b], c.n, s->n_pad, s->stream);
            if (e == hipSuccess && s->snap[b])
                e = launch_pad_f32(s->snap[b], c.n, s->n_pad, s->stream);
            if (e == hipSuccess && s->p64[b] && c.precision == UTCNB_PREC_FP64)
                e = launch_pad_f64(s->p64[b], c.n, s->n_pad, s->stream);
        }
        if (e != hipSuccess)
            rc = fail(UTCNB_EHIP, std::string("padding: ") + hipGetErrorString(e));
    }
    if (rc) {
        int keep = rc;
        release(s);
        delete s;
        return keep < 0 ? keep : UTCNB_ENOMEM;
    }
    *out = s;
    return UTCNB_OK;
}

int utcnb_destroy(utcnb_system* s)
{
    if (!s)
        return UTCNB_OK;
    (void)hipStreamSynchronize(s->stream);
    release(s);
    delete s;
    return UTCNB_OK;
}

int utcnb_set_state(utcnb_system* s, const void* pos4, const void* vel4, int dtype, int where)
{
    if (!s || !pos4)
        return fail(UTCNB_EINVAL, "utcnb_set_state: null handle or positions");
    if ((dtype != UTCNB_F32 && dtype != UTCNB_F64) || (where != UTCNB_HOST && where != UTCNB_DEVICE))
        return fail(UTCNB_EINVAL, "utcnb_set_state: bad dtype or location");
    std::vector<double> vdefault;
    int vdtype = dtype, vwhere = where;
    if (!vel4) {
        vdefault.assign(4 * (size_t)s->n, 0.0);
        for (int i = 0; i < s->n; ++i)
            vdefault[4 * (size_t)i + 3] = 1.0;
        vel4 = vdefault.data();
        vdtype = UTCNB_F64;
        vwhere = UTCNB_HOST;
    }
    s->cur = 0;
    int rc;
    if (is_f32_state(s)) {
        rc = upload4(s, pos4, dtype, where, s->p32[0]);
        if (!rc)
            rc = upload4(s, vel4, vdtype, vwhere, s->v32);
    } else {
        rc = upload4(s, pos4, dtype, where, s->p64[0]);
        if (!rc)
            rc = upload4(s, vel4, vdtype, vwhere, s->v64);
        if (!rc && s->cfg.precision == UTCNB_PREC_MIXED)
            UTCNB_HIP(launch_snapshot(s->p64[0], s->snap[0], s->n, s->stream));
    }
    if (rc)
        return rc;
    // host sources (the caller's, or the default velocities) may be reused or
    // freed once this returns: wait for the copies out of them
    if (where == UTCNB_HOST || !vdefault.empty())
        UTCNB_HIP(hipStreamSynchronize(s->stream));
    s->has_state = true;
    return UTCNB_OK;
}

int utcnb_get_state(utcnb_system* s, void* pos4, void* vel4, int dtype, int where)
{
    if (!s)
        return fail(UTCNB_EINVAL, "utcnb_get_state: null handle");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_get_state: no state set");
    if ((dtype != UTCNB_F32 && dtype != UTCNB_F64) || (where != UTCNB_HOST && where != UTCNB_DEVICE))
        return fail(UTCNB_EINVAL, "utcnb_get_state: bad dtype or location");
    int rc = UTCNB_OK;
    if (is_f32_state(s)) {
        if (pos4)
            rc = download4(s, s->p32[s->cur], pos4, dtype, where);
        if (!rc && vel4) {
            if (pos4 && where == UTCNB_HOST)
                UTCNB_HIP(hipStreamSynchronize(s->stream));  // staging reuse
            rc = download4(s, s->v32, vel4, dtype, where);
        }
    } else {
        if (pos4)
            rc = download4(s, cur_p64(s), pos4, dtype, where);
        if (!rc && vel4) {
            if (pos4)
                UTCNB_HIP(hipStreamSynchronize(s->stream));
            rc = download4(s, s->v64, vel4, dtype, where);
        }
    }
    if (rc)
        return rc;
    if (where == UTCNB_HOST)
        UTCNB_HIP(hipStreamSynchronize(s->stream));
    return UTCNB_OK;
}

// One step on the current buffers (no bookkeeping of timing).
static hipError_t step_once(utcnb_system* s)
{
    hipError_t e = hipSuccess;
    const int c = s->cur;
    switch (s->cfg.precision) {
    case UTCNB_PREC_REFERENCE:
        if (s->cfg.state_dtype == UTCNB_F32)
            e = launch_ref_step_f32(s->p32[c], s->p32[1 - c], s->v32, s->n, (float)s->cfg.dt,
                                    (float)s->cfg.eps2, (float)s->cfg.damping, 256, s->stream);
        else
            e = launch_ref_step_f64(s->p64[c], s->p64[1 - c], s->v64, s->n, s->cfg.dt,
                                    s->cfg.eps2, s->cfg.damping, 256, s->stream);
        s->cur = 1 - c;
        break;
    case UTCNB_PREC_FP32:
        e = launch_fp32_step(s->args, s->p32[c], s->p32[1 - c], s->v32, s->stream);
        s->cur = 1 - c;
        break;
    case UTCNB_PREC_MIXED:
        e = launch_mixed_step(s->args, s->snap[c], s->snap[1 - c], s->p64[0], s->v64, s->stream);
        s->cur = 1 - c;
        break;
    case UTCNB_PREC_FP64:
        e = launch_fp64_step(s->args, s->p64[c], s->p64[1 - c], s->v64, s->stream);
        s->cur = 1 - c;
        break;
    }
    return e;
}

int utcnb_step(utcnb_system* s, int nsteps)
{
    if (!s || nsteps < 0)
        return fail(UTCNB_EINVAL, "utcnb_step: null handle or negative step count");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_step: no state set");
    UTCNB_HIP(hipEventRecord(s->ev0, s->stream));
    for (int k = 0; k < nsteps; ++k) {
        hipError_t e = step_once(s);
        if (e != hipSuccess)
            return fail(UTCNB_EHIP, std::string("step kernel: ") + hipGetErrorString(e));
    }
    UTCNB_HIP(hipEventRecord(s->ev1, s->stream));
    s->last_launches = nsteps;
    s->timed = true;
    return UTCNB_OK;
}

int utcnb_sync(utcnb_system* s)
{
    if (!s)
        return fail(UTCNB_EINVAL, "utcnb_sync: null handle");
    UTCNB_HIP(hipStreamSynchronize(s->stream));
    return UTCNB_OK;
}

int utcnb_sym_chunks(const utcnb_system* s, int* chunks)
{
    if (!s || !chunks)
        return fail(UTCNB_EINVAL, "utcnb_sym_chunks: null argument");
    *chunks = s->args.chunks;
    return UTCNB_OK;
}

int utcnb_get_tile(const utcnb_system* s, int* tile)
{
    if (!s || !tile)
        return fail(UTCNB_EINVAL, "utcnb_get_tile: null argument");
    const bool tiled = s->cfg.precision == UTCNB_PREC_FP32 || s->cfg.precision == UTCNB_PREC_MIXED;
    *tile = tiled ? s->args.tile + 1 : 0;
    return UTCNB_OK;
}

int utcnb_last_kernel_ms(utcnb_system* s, double* ms, int* launches)
{
    if (!s || !ms)
        return fail(UTCNB_EINVAL, "utcnb_last_kernel_ms: null argument");
    if (!s->timed)
        return fail(UTCNB_ESTATE, "utcnb_last_kernel_ms: nothing timed yet");
    UTCNB_HIP(hipEventSynchronize(s->ev1));
    float f = 0.f;
    UTCNB_HIP(hipEventElapsedTime(&f, s->ev0, s->ev1));
    *ms = f;
    if (launches)
        *launches = s->last_launches;
    return UTCNB_OK;
}

// One-pass tile of a slice of `count` targets (0 for REFERENCE / FP64).
static int slice_tile_of(const utcnb_system* s, int count)
{
    if (s->cfg.precision != UTCNB_PREC_FP32 && s->cfg.precision != UTCNB_PREC_MIXED)
        return -1;
    return tile_get(s->args.tile).sym ? slice_tile_default(count, s->cu_count) : s->args.tile;
}

int utcnb_slice_tile(const utcnb_system* s, int count, int* tile)
{
    if (!s || !tile || count < 1)
        return fail(UTCNB_EINVAL, "utcnb_slice_tile: null argument or count < 1");
    *tile = slice_tile_of(s, count) + 1;
    return UTCNB_OK;
}

int utcnb_step_slice(utcnb_system* s, int start, int count)
{
    if (!s)
        return fail(UTCNB_EINVAL, "utcnb_step_slice: null handle");
    if (int rc = check_slice_args("utcnb_step_slice", s->n, start, count))
        return rc;
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_step_slice: no state set");
    const int c = s->cur;
    const int t = slice_tile_of(s, count);
    hipError_t e = hipSuccess;
    UTCNB_HIP(hipEventRecord(s->ev0, s->stream));
    switch (s->cfg.precision) {
    case UTCNB_PREC_REFERENCE:
        if (s->cfg.state_dtype == UTCNB_F32)
            e = launch_ref_step_slice_f32(s->p32[c], s->p32[1 - c] + start, s->v32 + start, s->n, start, count,
                                          (float)s->cfg.dt, (float)s->cfg.eps2, (float)s->cfg.damping, 256,
                                          s->stream);
        else
            e = launch_ref_step_slice_f64(s->p64[c], s->p64[1 - c] + start, s->v64 + start, s->n, start, count,
                                          s->cfg.dt, s->cfg.eps2, s->cfg.damping, 256, s->stream);
        break;
    case UTCNB_PREC_FP32:
        e = launch_fp32_step_slice(s->args, t, start, count, s->p32[c], s->p32[1 - c], s->v32, s->stream);
        break;
    case UTCNB_PREC_MIXED:
        e = launch_mixed_step_slice(s->args, t, start, count, s->snap[c], s->snap[1 - c], s->p64[0], s->v64,
                                    s->stream);
        break;
    case UTCNB_PREC_FP64:
        e = launch_fp64_step_slice(s->args, start, count, s->p64[c], s->p64[1 - c], s->v64, s->stream);
        break;
    }
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_step_slice: ") + hipGetErrorString(e));
    UTCNB_HIP(hipEventRecord(s->ev1, s->stream));
    s->last_launches = 1;
    s->timed = true;
    return UTCNB_OK;
}

int utcnb_commit_step(utcnb_system* s)
{
    if (!s)
        return fail(UTCNB_EINVAL, "utcnb_commit_step: null handle");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_commit_step: no state set");
    s->cur = 1 - s->cur;
    return UTCNB_OK;
}

int utcnb_position_buffer(utcnb_system* s, int which, void** pos4, int* dtype)
{
    if (!s || !pos4 || (which != 0 && which != 1))
        return fail(UTCNB_EINVAL, "utcnb_position_buffer: null argument or which not 0/1");
    const int b = which == 0 ? s->cur : 1 - s->cur;
    int dt = UTCNB_F32;
    void* p = nullptr;
    switch (s->cfg.precision) {
    case UTCNB_PREC_REFERENCE:
        if (s->cfg.state_dtype == UTCNB_F32) {
            p = s->p32[b];
        } else {
            p = s->p64[b];
            dt = UTCNB_F64;
        }
        break;
    case UTCNB_PREC_FP32: p = s->p32[b]; break;
    case UTCNB_PREC_MIXED: p = s->snap[b]; break;
    case UTCNB_PREC_FP64:
        p = s->p64[b];
        dt = UTCNB_F64;
        break;
    }
    *pos4 = p;
    if (dtype)
        *dtype = dt;
    return UTCNB_OK;
}

// ---- shares of a symmetric-tile system (wave-block partition) -------------------
struct ShareGeom {
    int NW, w0, w1, start, count, Rs;
};

static int share_geom(const utcnb_system* s, int shares, int share, const char* who, ShareGeom* g)
{
    if (s->cfg.precision != UTCNB_PREC_MIXED || !tile_get(s->args.tile).sym)
        return fail(UTCNB_EINVAL, std::string(who) + ": needs a MIXED system with a symmetric tile "
                                                     "(utcnb_step_slice shares any other system)");
    if (!s->jpart)
        return fail(UTCNB_EINVAL, std::string(who) + ": the system evaluates its symmetric steps in chunks "
                                                     "(no whole-system reaction buffer); use utcnb_step_slice");
    const Tile& t = tile_get(s->args.tile);
    const int B = 64 * t.k;
    g->NW = s->n_pad / B;
    if (shares < 1 || shares > g->NW || share < 0 || share >= shares)
        return fail(UTCNB_EINVAL, std::string(who) + ": need 1 <= shares <= wave blocks (" +
                                      std::to_string(g->NW) + ") and 0 <= share < shares");
    sym_share_blocks(g->NW, shares, share, &g->w0, &g->w1);
    g->start = std::min(g->w0 * B, s->n);
    g->count = std::min(g->w1 * B, s->n) - g->start;
    // one split count for every share of the partition (from the largest
    // share), so the shares' target-side partials share one layout
    g->Rs = sym_share_splits((g->NW + shares - 1) / shares, t.s, t.k, g->NW);
    return UTCNB_OK;
}

int utcnb_share_range(const utcnb_system* s, int shares, int share, int* start, int* count)
{
    if (!s || !start || !count)
        return fail(UTCNB_EINVAL, "utcnb_share_range: null argument");
    ShareGeom g;
    if (int rc = share_geom(s, shares, share, "utcnb_share_range", &g))
        return rc;
    *start = g.start;
    *count = g.count;
    return UTCNB_OK;
}

int utcnb_share_pairs(utcnb_system* s, int shares, int share, double* react3)
{
    if (!s || !react3)
        return fail(UTCNB_EINVAL, "utcnb_share_pairs: null argument");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_share_pairs: no state set");
    ShareGeom g;
    if (int rc = share_geom(s, shares, share, "utcnb_share_pairs", &g))
        return rc;
    const size_t need = (size_t)g.Rs * (size_t)s->n_pad * 3;
    if (need > s->share_ipart_floats) {
        if (need * sizeof(float) > s->partial_budget)
            return fail(UTCNB_EINVAL, "utcnb_share_pairs: the share's partial sums exceed the device budget");
        dfree(s->share_ipart);
        s->share_ipart_floats = 0;
        if (int rc = dalloc(&s->share_ipart, need))
            return rc;
        s->share_ipart_floats = need;
    }
    UTCNB_HIP(hipEventRecord(s->ev0, s->stream));
    hipError_t e = launch_sym_share_pairs(tile_get(s->args.tile), s->args.newton, s->snap[s->cur], s->n, s->n_pad,
                                          s->args.eps2f, shares, share, g.Rs, s->share_ipart, true, s->jpart,
                                          false, react3, false, s->stream);
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_share_pairs: ") + hipGetErrorString(e));
    UTCNB_HIP(hipEventRecord(s->ev1, s->stream));
    s->last_launches = 1;
    s->timed = true;
    return UTCNB_OK;
}

int utcnb_share_finish(utcnb_system* s, int shares, int share, const double* react_sum3)
{
    if (!s)
        return fail(UTCNB_EINVAL, "utcnb_share_finish: null handle");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_share_finish: no state set");
    ShareGeom g;
    if (int rc = share_geom(s, shares, share, "utcnb_share_finish", &g))
        return rc;
    if (g.count <= 0)
        return UTCNB_OK;  // a share of padding bodies only: nothing to finish
    if (!react_sum3)
        return fail(UTCNB_EINVAL, "utcnb_share_finish: null react_sum3");
    if (!s->share_ipart || (size_t)g.Rs * (size_t)s->n_pad * 3 > s->share_ipart_floats)
        return fail(UTCNB_ESTATE, "utcnb_share_finish: utcnb_share_pairs of this share has not run");
    UTCNB_HIP(hipEventRecord(s->ev0, s->stream));
    hipError_t e = launch_sym_share_finish_mixed(s->args, shares, share, g.Rs, g.start, g.count, s->share_ipart,
                                                 react_sum3, s->snap[1 - s->cur], s->p64[0], s->v64, s->stream);
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_share_finish: ") + hipGetErrorString(e));
    UTCNB_HIP(hipEventRecord(s->ev1, s->stream));
    s->last_launches = 1;
    s->timed = true;
    return UTCNB_OK;
}

int utcnb_accel_slice(utcnb_system* s, int start, int count, double* acc3)
{
    if (!s || !acc3)
        return fail(UTCNB_EINVAL, "utcnb_accel_slice: null argument");
    if (int rc = check_slice_args("utcnb_accel_slice", s->n, start, count))
        return rc;
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_accel_slice: no state set");
    const int t = slice_tile_of(s, count);
    hipError_t e = hipSuccess;
    switch (s->cfg.precision) {
    case UTCNB_PREC_REFERENCE:
        if (s->cfg.state_dtype == UTCNB_F32)
            e = launch_ref_accel_slice_f32(s->p32[s->cur], s->n, start, count, (float)s->cfg.eps2, s->scratch,
                                           s->stream);
        else
            e = launch_ref_accel_slice_f64(s->p64[s->cur], s->n, start, count, s->cfg.eps2, s->scratch, s->stream);
        break;
    case UTCNB_PREC_FP32:
        e = launch_fp32_accel_slice(s->args, t, start, count, s->p32[s->cur], false, s->scratch, s->stream);
        break;
    case UTCNB_PREC_MIXED:
        e = launch_fp32_accel_slice(s->args, t, start, count, s->snap[s->cur], true, s->scratch, s->stream);
        break;
    case UTCNB_PREC_FP64:
        e = launch_fp64_accel_slice(s->args, start, count, s->p64[s->cur], s->scratch, s->stream);
        break;
    }
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_accel_slice: ") + hipGetErrorString(e));
    UTCNB_HIP(hipMemcpyAsync(acc3, s->scratch, sizeof(double) * 3 * (size_t)count, hipMemcpyDeviceToHost,
                             s->stream));
    UTCNB_HIP(hipStreamSynchronize(s->stream));
    return UTCNB_OK;
}

int utcnb_accel(utcnb_system* s, double* acc3)
{
    if (!s || !acc3)
        return fail(UTCNB_EINVAL, "utcnb_accel: null argument");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_accel: no state set");
    hipError_t e = hipSuccess;
    switch (s->cfg.precision) {
    case UTCNB_PREC_REFERENCE:
        if (s->cfg.state_dtype == UTCNB_F32)
            e = launch_ref_accel_f32(s->p32[s->cur], s->n, (float)s->cfg.eps2, s->scratch,
                                     s->stream);
        else
            e = launch_ref_accel_f64(s->p64[s->cur], s->n, s->cfg.eps2, s->scratch, s->stream);
        break;
    case UTCNB_PREC_FP32:
        e = launch_fp32_accel(s->args, s->p32[s->cur], false, s->scratch, s->stream);
        break;
    case UTCNB_PREC_MIXED:
        e = launch_fp32_accel(s->args, s->snap[s->cur], true, s->scratch, s->stream);
        break;
    case UTCNB_PREC_FP64:
        e = launch_fp64_accel(s->args, s->p64[s->cur], s->scratch, s->stream);
        break;
    }
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("field kernel: ") + hipGetErrorString(e));
    UTCNB_HIP(hipMemcpyAsync(acc3, s->scratch, sizeof(double) * 3 * (size_t)s->n,
                             hipMemcpyDeviceToHost, s->stream));
    UTCNB_HIP(hipStreamSynchronize(s->stream));
    return UTCNB_OK;
}


// ---- adjoint -------------------------------------------------------------------
static int adjoint_workspace(utcnb_system* s, int nck, int seg)
{
    auto& a = s->adj;
    const size_t n = (size_t)s->n, np = (size_t)s->n_pad;
    const bool f32 = s->cfg.precision == UTCNB_PREC_MIXED;
    if (a.xbar && a.nck >= nck && a.seg >= seg)
        return UTCNB_OK;
    release_adjoint(s);
    int rc = dalloc(&a.ckpt, 2 * n * (size_t)nck) || dalloc(&a.segpos, n * (size_t)seg) ||
             (nck > 1 && dalloc(&a.segpos2, n * (size_t)seg)) ||
             dalloc(&a.wpos[0], np) || dalloc(&a.wpos[1], np) || dalloc(&a.wvel, n) ||
             dalloc(&a.xbar, 3 * n) || dalloc(&a.vbar, 3 * n) || dalloc(&a.mbar, n) ||
             dalloc(&a.tgt, 3 * n) || dalloc(&a.part, (n + 255) / 256 + 1);
    if (!rc && f32)
        rc = dalloc(&a.wsnap[0], np) || dalloc(&a.wsnap[1], np);
    const int Rv = sym_vjp_splits((int)np, s->cu_count);
    const Tile& ft = tile_get(s->args.tile);
    if (!rc && f32 && (ft.sym || ft.cap > 0) && sym_vjp_applies((int)np) &&
        (sym_vjp_jpart_float4s((int)np) + sym_vjp_ipart_float4s((int)np, Rv)) * sizeof(float4) <=
            s->partial_budget) {
        // a symmetric force tile (or a small-system one, whose sizes the
        // symmetric VJP also serves best) evaluates the adjoint pair sum
        // symmetrically
        const int R = Rv;
        rc = dalloc(&a.vjp_ipart, sym_vjp_ipart_float4s((int)np, R)) ||
             dalloc(&a.vjp_jpart, sym_vjp_jpart_float4s((int)np));
        a.vjp_r = R;
    }
    if (!rc) {
        const size_t rb = f32 ? sizeof(float4) : sizeof(double4);
        if (hipMalloc(&a.jrec, 2 * np * rb) != hipSuccess || hipMalloc(&a.fsrc, np * rb) != hipSuccess)
            rc = 1;
    }
    if (!rc && nck > 1) {
        if (hipStreamCreateWithFlags(&a.rstream, hipStreamNonBlocking) != hipSuccess)
            rc = 1;
        for (hipEvent_t* ev : {&a.evfwd, &a.evrec[0], &a.evrec[1], &a.evvjp})
            if (!rc && hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess)
                rc = 1;
    }
    if (rc) {
        release_adjoint(s);
        return fail(UTCNB_ENOMEM, "utcnb_adjoint: workspace allocation failed");
    }
    a.nck = nck;
    a.seg = seg;
    hipError_t e = launch_jrec_pad(f32, a.jrec, s->n, s->n_pad, s->stream);
    if (e == hipSuccess && f32)
        e = launch_pad_f32((float4*)a.fsrc, s->n, s->n_pad, s->stream);
    if (e == hipSuccess && f32)
        e = launch_pad_f32(a.wsnap[0], s->n, s->n_pad, s->stream);
    if (e == hipSuccess && f32)
        e = launch_pad_f32(a.wsnap[1], s->n, s->n_pad, s->stream);
    if (e == hipSuccess && !f32)
        e = launch_pad_f64((double4*)a.fsrc, s->n, s->n_pad, s->stream);
    if (e == hipSuccess)
        e = launch_pad_f64(a.wpos[0], s->n, s->n_pad, s->stream);
    if (e == hipSuccess)
        e = launch_pad_f64(a.wpos[1], s->n, s->n_pad, s->stream);
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_adjoint: padding: ") + hipGetErrorString(e));
    return UTCNB_OK;
}

// One forward step of the adjoint's work state (wpos/wvel, with wsnap for MIXED).
static hipError_t work_step(utcnb_system* s, int& wc, hipStream_t st)
{
    auto& a = s->adj;
    hipError_t e;
    if (s->cfg.precision == UTCNB_PREC_MIXED) {
        e = launch_mixed_step(s->args, a.wsnap[wc], a.wsnap[1 - wc], a.wpos[0], a.wvel, st);
    } else {
        e = launch_fp64_step(s->args, a.wpos[wc], a.wpos[1 - wc], a.wvel, st);
    }
    wc = 1 - wc;
    return e;
}

int utcnb_adjoint(utcnb_system* s, int nsteps, int ckpt_every, const double* target3,
                             double* loss, double* dx0, double* dv0, double* dm)
{
    if (!s || !target3 || !loss || !dx0 || !dv0 || !dm)
        return fail(UTCNB_EINVAL, "utcnb_adjoint: null argument");
    if (s->cfg.precision != UTCNB_PREC_MIXED && s->cfg.precision != UTCNB_PREC_FP64)
        return fail(UTCNB_EINVAL, "utcnb_adjoint: precision must be MIXED or FP64");
    if (nsteps < 1 || ckpt_every < 1)
        return fail(UTCNB_EINVAL, "utcnb_adjoint: nsteps and ckpt_every must be >= 1");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_adjoint: no state set");
    const int k = ckpt_every < nsteps ? ckpt_every : nsteps;
    const int nck = (nsteps + k - 1) / k;
    int rc = adjoint_workspace(s, nck, k);
    if (rc)
        return rc;
    auto& a = s->adj;
    const bool mixed = s->cfg.precision == UTCNB_PREC_MIXED;
    const size_t n = (size_t)s->n;
    const size_t sb = n * sizeof(double4);
    const bool refkick = s->cfg.kick == UTCNB_KICK_REFERENCE;
    hipStream_t st = s->stream;
    int launches = 0;
#define ADJ_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(UTCNB_EHIP, std::string("utcnb_adjoint: ") + #expr + ": " +     \
                                        hipGetErrorString(e_));                          \
    } while (0)

    ADJ_CHECK(hipMemcpyAsync(a.tgt, target3, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
    ADJ_CHECK(hipEventRecord(s->ev0, st));
    // forward sweep with checkpoints of (x, v) at steps 0, k, 2k, ...; the
    // positions of the last segment go straight to the segment buffer, which
    // the reverse sweep would otherwise fill by recomputing that segment (the
    // recompute runs the same kernels on the same state, so the positions are
    // the same bits)
    const int slast = (nck - 1) * k;
    // segment c's positions live in segbuf(c): two buffers, so that the
    // recompute of segment c - 1 can run beside the VJP steps of segment c
    auto segbuf = [&](int c) { return (c % 2 == 0) ? a.segpos : a.segpos2; };
    for (int t = 0; t < nsteps; ++t) {
        if (t % k == 0) {
            double4* ck = a.ckpt + 2 * n * (size_t)(t / k);
            ADJ_CHECK(hipMemcpyAsync(ck, cur_p64(s), sb, hipMemcpyDeviceToDevice, st));
            ADJ_CHECK(hipMemcpyAsync(ck + n, s->v64, sb, hipMemcpyDeviceToDevice, st));
        }
        if (t >= slast)
            ADJ_CHECK(hipMemcpyAsync(segbuf(nck - 1) + n * (size_t)(t - slast), cur_p64(s), sb,
                                     hipMemcpyDeviceToDevice, st));
        ADJ_CHECK(step_once(s));
        ++launches;
    }
    // loss and its seed at the final state
    double* dloss = a.part + (n + 255) / 256;
    ADJ_CHECK(launch_loss_seed(cur_p64(s), a.tgt, s->n, a.xbar, a.vbar, a.mbar, a.part, dloss, st));
    // Reverse sweep, one checkpoint segment at a time.  Every segment but the
    // last is recomputed from its checkpoint on a stream of its own (rs), one
    // segment ahead of the VJP steps on st: the VJP of segment c waits for
    // its recompute (evrec), and the recompute of segment c - 2, which reuses
    // segment c's buffer, waits for that VJP (evvjp).  The reference kick's
    // field evaluation in the VJP loop shares the force kernels' partial sums
    // with the recompute, so that kick runs both on st, in the same order.
    const bool overlap = nck > 1 && !refkick;
    hipStream_t rs = overlap ? a.rstream : st;
    auto recompute = [&](int c) -> hipError_t {
        const int s0 = c * k, s1 = (s0 + k < nsteps) ? s0 + k : nsteps;
        const double4* ck = a.ckpt + 2 * n * (size_t)c;
        double4* buf = segbuf(c);
        hipError_t e = hipMemcpyAsync(a.wpos[0], ck, sb, hipMemcpyDeviceToDevice, rs);
        if (e == hipSuccess)
            e = hipMemcpyAsync(a.wvel, ck + n, sb, hipMemcpyDeviceToDevice, rs);
        if (e == hipSuccess && mixed)
            e = launch_snapshot(a.wpos[0], a.wsnap[0], s->n, rs);
        int wc = 0;
        for (int t = s0; t < s1 && e == hipSuccess; ++t) {
            const double4* cur = mixed ? a.wpos[0] : a.wpos[wc];
            e = hipMemcpyAsync(buf + n * (size_t)(t - s0), cur, sb, hipMemcpyDeviceToDevice, rs);
            if (e == hipSuccess && t + 1 < s1) {
                e = work_step(s, wc, rs);
                ++launches;
            }
        }
        if (e == hipSuccess && overlap)
            e = hipEventRecord(a.evrec[c % 2], rs);
        return e;
    };
    if (overlap) {
        ADJ_CHECK(hipEventRecord(a.evfwd, st));  // checkpoints and force partials free
        ADJ_CHECK(hipStreamWaitEvent(rs, a.evfwd, 0));
    }
    if (nck > 1)
        ADJ_CHECK(recompute(nck - 2));
    for (int c = nck - 1; c >= 0; --c) {
        const int s0 = c * k;
        const int s1 = (s0 + k < nsteps) ? s0 + k : nsteps;
        if (overlap && c < nck - 1)
            ADJ_CHECK(hipStreamWaitEvent(st, a.evrec[c % 2], 0));
        for (int t = s1 - 1; t >= s0; --t) {
            const double4* xs = segbuf(c) + n * (size_t)(t - s0);
            const double* acc = nullptr;
            if (refkick) {
                if (mixed) {
                    ADJ_CHECK(launch_snapshot(xs, (float4*)a.fsrc, s->n, st));
                    ADJ_CHECK(launch_fp32_accel(s->args, (const float4*)a.fsrc, true, s->scratch, st));
                } else {
                    ADJ_CHECK(hipMemcpyAsync(a.fsrc, xs, sb, hipMemcpyDeviceToDevice, st));
                    ADJ_CHECK(launch_fp64_accel(s->args, (const double4*)a.fsrc, s->scratch, st));
                }
                acc = s->scratch;
                ++launches;
            }
            ADJ_CHECK(launch_adjoint_pre(mixed, xs, acc, s->n, s->cfg.dt, s->cfg.damping,
                                         s->cfg.kick, a.xbar, a.vbar, a.mbar, a.jrec, st));
            if (mixed && a.vjp_r > 0)
                ADJ_CHECK(launch_sym_vjp((const float4*)a.jrec, s->n, s->n_pad, s->args.eps2f, a.vjp_r,
                                         a.vjp_ipart, a.vjp_jpart, a.xbar, a.mbar, st));
            else if (mixed)
                ADJ_CHECK(launch_vjp_f32((const float4*)a.jrec, s->n, s->n_pad, s->cfg.eps2,
                                         a.xbar, a.mbar, s->cu_count, st));
            else
                ADJ_CHECK(launch_vjp_f64((const double4*)a.jrec, s->n, s->n_pad, s->cfg.eps2,
                                         a.xbar, a.mbar, s->cu_count, st));
            ++launches;
        }
        if (c >= 2) {
            if (overlap) {
                ADJ_CHECK(hipEventRecord(a.evvjp, st));  // segment c's buffer is free again
                ADJ_CHECK(hipStreamWaitEvent(rs, a.evvjp, 0));
            }
            ADJ_CHECK(recompute(c - 2));
        }
    }
    ADJ_CHECK(hipEventRecord(s->ev1, st));
    s->last_launches = launches;
    s->timed = true;
    ADJ_CHECK(hipMemcpyAsync(dx0, a.xbar, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    ADJ_CHECK(hipMemcpyAsync(dv0, a.vbar, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    ADJ_CHECK(hipMemcpyAsync(dm, a.mbar, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ADJ_CHECK(hipMemcpyAsync(loss, dloss, sizeof(double), hipMemcpyDeviceToHost, st));
    ADJ_CHECK(hipStreamSynchronize(st));
#undef ADJ_CHECK
    return UTCNB_OK;
}

int utcnb_diagnostics(utcnb_system* s, double* kinetic, double* potential, double* momentum3)
{
    if (!s)
        return fail(UTCNB_EINVAL, "utcnb_diagnostics: null handle");
    if (!s->has_state)
        return fail(UTCNB_ESTATE, "utcnb_diagnostics: no state set");
    const double4* p;
    const double4* v;
    if (is_f32_state(s)) {
        double4* tp = (double4*)s->stage;
        double4* tv = tp + s->n;
        UTCNB_HIP(launch_f32_to_f64(s->p32[s->cur], tp, s->n, s->stream));
        UTCNB_HIP(launch_f32_to_f64(s->v32, tv, s->n, s->stream));
        p = tp;
        v = tv;
    } else {
        p = cur_p64(s);
        v = s->v64;
    }
    UTCNB_HIP(launch_energy(p, v, s->n, s->cfg.eps2, s->scratch, s->stream));
    double out[5];
    UTCNB_HIP(hipMemcpyAsync(out, s->scratch, sizeof(out), hipMemcpyDeviceToHost, s->stream));
    UTCNB_HIP(hipStreamSynchronize(s->stream));
    if (kinetic)
        *kinetic = out[0];
    if (potential)
        *potential = out[1];
    if (momentum3) {
        momentum3[0] = out[2];
        momentum3[1] = out[3];
        momentum3[2] = out[4];
    }
    return UTCNB_OK;
}

}  // extern "C"
