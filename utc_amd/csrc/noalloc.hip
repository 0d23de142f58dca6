// noalloc.hip -- utcnb_accel_noalloc: the field of n bodies given on the
// device, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The work memory the field kernels need beyond the caller's input and
// output -- the padded FP32 source array and, for the symmetric tiles, their
// target-side and reaction partial sums -- lives in `scratch`, a file-scope
// device array of kScratchBytes.  Each device the library runs on holds its
// own copy until the process ends.  Host code never passes the array's name
// to a kernel (in host code it is only a stand-in for the symbol): the first
// call on a device, with that device current, resolves the device address
// with hipGetSymbolAddress and keeps it per device; every later use goes
// through that address.
//
// Nothing in the scratch needs to start at zero: the source copy is written
// in full by copy_pad_kernel, and every partial-sum record the symmetric
// kernels read is written by the same launch before it is read.  All calls on
// one device share its scratch, so calls that overlap on one device (two
// streams, two threads) race and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "internal.h"

namespace utcnb {

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
    void* ptr = nullptr;  // device address of `scratch` on this device
    int cus = 0;
    bool ok = false;
};
DeviceScratch g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

// Resolves (once per device) the address of this device's scratch array.
int scratch_of_current_device(DeviceScratch* out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return fail(UTCNB_ENODEV, "utcnb_accel_noalloc: no HIP device available");
    if (dev < 0 || dev >= kMaxDevices)
        return fail(UTCNB_ENODEV, "utcnb_accel_noalloc: device index beyond the supported 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    DeviceScratch& d = g_scratch[dev];
    if (!d.ok) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
            return fail(UTCNB_ENODEV, "utcnb_accel_noalloc: hipGetDeviceProperties failed");
        if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
            return fail(UTCNB_ENODEV, std::string("utcnb_accel_noalloc: device is ") + prop.gcnArchName +
                                          ", this library is built for gfx950 only");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(UTCNB_EHIP, std::string("utcnb_accel_noalloc: hipGetSymbolAddress: ") +
                                        hipGetErrorString(e));
        d.ptr = p;
        d.cus = prop.multiProcessorCount;
        d.ok = true;
    }
    *out = d;
    return UTCNB_OK;
}

// dst[i] = src[i] for i < n, a far-away zero-mass padding body for n <= i < n_pad.
__global__ void copy_pad_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int n, int n_pad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
    else if (i < n_pad)
        dst[i] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

}  // namespace utcnb

using namespace utcnb;

extern "C" int utcnb_accel_noalloc(const float* pos4, int n, double eps2, int precision, int tile, int flags,
                                   double* acc3, void* stream)
{
    if (!pos4 || !acc3 || n < 1)
        return fail(UTCNB_EINVAL, "utcnb_accel_noalloc: null pointer or n < 1");
    if (precision != UTCNB_PREC_FP32 && precision != UTCNB_PREC_MIXED)
        return fail(UTCNB_EINVAL, "utcnb_accel_noalloc: precision must be FP32 or MIXED");
    if (!(eps2 > 0.0) || !(eps2 < 1e300))
        return fail(UTCNB_EINVAL, "utcnb_accel_noalloc: eps2 must be finite and > 0");
    if (tile < 0 || tile > tile_count())
        return fail(UTCNB_EINVAL, "utcnb_accel_noalloc: tile out of range");
    DeviceScratch ds;
    int rc = scratch_of_current_device(&ds);
    if (rc)
        return rc;
    const int64_t n_pad = pad_j(n);
    const size_t src_bytes = round256((size_t)n_pad * sizeof(float4));
    if (src_bytes > kScratchBytes)
        return fail(UTCNB_EINVAL, "utcnb_accel_noalloc: n = " + std::to_string(n) +
                                      " needs more than the 64 MiB scratch");
    // bytes of scratch tile t needs, and where its partial sums go
    size_t ipart_off = 0, jpart_off = 0;
    auto need_of = [&](int t) {
        const Tile& tl = tile_get(t);
        size_t need = src_bytes;
        if (tl.sym) {
            ipart_off = need;
            need += round256(sym_ipart_floats((int)n_pad, tl.k, sym_tile_r(tl, (int)n_pad)) * sizeof(float));
            jpart_off = need;
            need += round256(sym_jpart_floats((int)n_pad, tl.k, tl.rs) * sizeof(float));
        }
        return need;
    };
    int t = tile > 0 ? tile - 1 : tile_default(n, ds.cus, kScratchBytes - src_bytes);
    if (tile == 0 && need_of(t) > kScratchBytes)
        t = tile_default(n, ds.cus, 0);  // the best one-pass tile: no partial sums
    const Tile& tl = tile_get(t);
    if (tl.sym && n_pad % (128 * tl.k) != 0)
        return fail(UTCNB_EINVAL, "utcnb_accel_noalloc: symmetric tile needs N padded to 128*K");
    if (tl.cap > 0 && n_pad > tl.cap)
        return fail(UTCNB_EINVAL, std::string("utcnb_accel_noalloc: tile ") + tl.name + " holds at most " +
                                      std::to_string(tl.cap) + " bodies (N padded to 1024)");
    const size_t need = need_of(t);
    if (need > kScratchBytes)
        return fail(UTCNB_EINVAL, "utcnb_accel_noalloc: n = " + std::to_string(n) + " with tile " + tl.name +
                                      " needs " + std::to_string(need) +
                                      " bytes of work memory, above the 64 MiB scratch");
    unsigned char* base = (unsigned char*)ds.ptr;
    float4* src = (float4*)base;
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(copy_pad_kernel, dim3((unsigned)((n_pad + 255) / 256)), dim3(256), 0, st,
                       (const float4*)pos4, src, n, (int)n_pad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_accel_noalloc: copy: ") + hipGetErrorString(e));
    StepArgs a{};
    a.n = n;
    a.n_pad = (int)n_pad;
    a.eps2 = eps2;
    a.eps2f = (float)eps2;
    a.newton = (flags & UTCNB_FLAG_NO_NEWTON) == 0;
    a.tile = t;
    a.ipart = tl.sym ? (float*)(base + ipart_off) : nullptr;
    a.jpart = tl.sym ? (float*)(base + jpart_off) : nullptr;
    e = launch_fp32_accel(a, src, precision == UTCNB_PREC_MIXED, acc3, st);
    if (e != hipSuccess)
        return fail(UTCNB_EHIP, std::string("utcnb_accel_noalloc: field kernel: ") + hipGetErrorString(e));
    return UTCNB_OK;
}
