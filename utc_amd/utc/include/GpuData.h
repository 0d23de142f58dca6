// GpuData.h -- a host buffer paired with a device buffer, with a record of
// which side holds the latest contents.
//
// Same public surface and status rules as the reference's
// include/gpu/GpuData.h:17-354 and include/gpu/GpuData.inc:1-158:
//   pageable: new[] host + device allocation; pinned: page-locked host
//   (hipHostMalloc) + device; unified: one managed allocation for both.
//   initH/putH mark the host side newer, initD/putD the device side;
//   sync() copies toward the older side; fetch() returns the newest contents.
// Errors of HIP calls are printed and execution continues, as in the
// reference (include/gpu/helper_cuda.h:419-446).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>

namespace iUtc {

enum class MemType { unknown = -1, pageable, pinned, unified };
enum class MemStatus { unknown = -1, synced, host, device };

inline void utcHipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        std::fprintf(stderr, "HIP error %s at %s\n", hipGetErrorString(e), what);
}

template <typename T>
class GpuData {
public:
    typedef T _Type;

    explicit GpuData(MemType memtype = MemType::pageable) : m_memtype(memtype) {}
    GpuData(unsigned long size, MemType memtype = MemType::pageable) { alloc(size, 1, 1, memtype, 1); }
    GpuData(unsigned long sx, unsigned long sy, MemType memtype = MemType::pageable)
    {
        alloc(sx, sy, 1, memtype, 2);
    }
    GpuData(unsigned long sx, unsigned long sy, unsigned long sz, MemType memtype = MemType::pageable)
    {
        alloc(sx, sy, sz, memtype, 3);
    }
    GpuData(const GpuData&) = delete;
    GpuData& operator=(const GpuData&) = delete;
    ~GpuData() { cleanMem(); }

    int initMem(unsigned long sx, MemType memtype = MemType::pageable) { return initMem(sx, 1, 1, memtype); }
    int initMem(unsigned long sx, unsigned long sy, MemType memtype = MemType::pageable)
    {
        return initMem(sx, sy, 1, memtype);
    }
    int initMem(unsigned long sx, unsigned long sy, unsigned long sz, MemType memtype = MemType::pageable)
    {
        if (!m_init)
            alloc(sx, sy, sz, memtype, 3);
        return 0;
    }

    void cleanMem()
    {
        if (!m_init)
            return;
        switch (m_memtype) {
        case MemType::pageable:
            delete[] m_host;
            utcHipCheck(hipFree(m_dev), "GpuData free");
            break;
        case MemType::pinned:
            utcHipCheck(hipHostFree(m_host), "GpuData free host");
            utcHipCheck(hipFree(m_dev), "GpuData free");
            break;
        case MemType::unified:
            utcHipCheck(hipFree(m_dev), "GpuData free managed");
            break;
        default:
            break;
        }
        m_host = m_dev = nullptr;
        m_init = false;
    }

    // ---- addresses -------------------------------------------------------------
    T* get() { return newest_is_host() ? m_host : m_dev; }
    T* get(bool modified)
    {
        T* p = get();
        if (modified)
            m_status = (p == m_host && m_memtype != MemType::unified) ? MemStatus::host : MemStatus::device;
        return p;
    }
    T* getH() { return m_host; }
    T* getD() { return m_dev; }
    T* getH(bool modified)
    {
        if (modified)
            m_status = MemStatus::host;
        return m_host;
    }
    T* getD(bool modified)
    {
        if (modified)
            m_status = MemStatus::device;
        return m_dev;
    }
    T* getH(unsigned long off) { return m_host + off; }
    T* getD(unsigned long off) { return m_dev + off; }
    void setH() { m_status = MemStatus::host; }
    void setD() { m_status = MemStatus::device; }
    MemStatus getStatus() const { return m_status; }
    MemType getMemType() const { return m_memtype; }
    unsigned long getSize() const { return m_size[0] * m_size[1] * m_size[2]; }
    unsigned long getBSize() const { return m_bytes; }

    // ---- synchronisation ---------------------------------------------------------
    void syncH() { copy_to_device(); }
    void syncD() { copy_to_host(); }
    void sync()
    {
        if (m_memtype == MemType::unified) {
            utcHipCheck(hipDeviceSynchronize(), "GpuData sync");
            m_status = MemStatus::synced;
            return;
        }
        if (m_status == MemStatus::host)
            copy_to_device();
        else if (m_status == MemStatus::device)
            copy_to_host();
    }

    // ---- contents ----------------------------------------------------------------
    void initH(int value = 0)
    {
        std::memset(m_host, value, m_bytes);
        m_status = MemStatus::host;
    }
    void initD(int value = 0)
    {
        utcHipCheck(hipMemset(m_dev, value, m_bytes), "GpuData initD");
        m_status = MemStatus::device;
    }
    void init(int value = 0)
    {
        std::memset(m_host, value, m_bytes);
        if (m_memtype != MemType::unified)
            utcHipCheck(hipMemset(m_dev, value, m_bytes), "GpuData init");
        m_status = MemStatus::synced;
    }
    void initH(const T* src)
    {
        std::memcpy(m_host, src, m_bytes);
        m_status = MemStatus::host;
    }
    void initD(const T* src)
    {
        utcHipCheck(hipMemcpy(m_dev, src, m_bytes, hipMemcpyDefault), "GpuData initD");
        m_status = MemStatus::device;
    }
    void init(const T* src)
    {
        std::memcpy(m_host, src, m_bytes);
        if (m_memtype != MemType::unified)
            utcHipCheck(hipMemcpy(m_dev, src, m_bytes, hipMemcpyDefault), "GpuData init");
        m_status = MemStatus::synced;
    }
    void putH(const T* src) { initH(src); }
    void putD(const T* src) { initD(src); }
    void put(const T* src) { init(src); }
    void fetchH(T* dst) { std::memcpy(dst, m_host, m_bytes); }
    void fetchD(T* dst) { utcHipCheck(hipMemcpy(dst, m_dev, m_bytes, hipMemcpyDefault), "GpuData fetchD"); }
    void fetch(T* dst)
    {
        if (!newest_is_host()) {
            utcHipCheck(hipMemcpy(m_host, m_dev, m_bytes, hipMemcpyDefault), "GpuData fetch");
            m_status = MemStatus::synced;
        }
        std::memcpy(dst, m_host, m_bytes);
    }
    T at(unsigned long i) const { return m_host[i]; }
    void assign(unsigned long i, const T v)
    {
        m_host[i] = v;
        m_status = MemStatus::host;
    }

private:
    bool newest_is_host() const
    {
        return m_memtype == MemType::unified || m_status == MemStatus::host ||
               m_status == MemStatus::synced;
    }
    void copy_to_device()
    {
        if (m_memtype != MemType::unified)
            utcHipCheck(hipMemcpy(m_dev, m_host, m_bytes, hipMemcpyHostToDevice), "GpuData H2D");
        m_status = MemStatus::synced;
    }
    void copy_to_host()
    {
        if (m_memtype != MemType::unified)
            utcHipCheck(hipMemcpy(m_host, m_dev, m_bytes, hipMemcpyDeviceToHost), "GpuData D2H");
        else
            utcHipCheck(hipDeviceSynchronize(), "GpuData sync");
        m_status = MemStatus::synced;
    }
    void alloc(unsigned long sx, unsigned long sy, unsigned long sz, MemType memtype, int dim)
    {
        m_dim = dim;
        m_size[0] = sx;
        m_size[1] = sy;
        m_size[2] = sz;
        m_bytes = sx * sy * sz * sizeof(T);
        m_memtype = memtype;
        switch (memtype) {
        case MemType::pageable:
            m_host = new T[sx * sy * sz];
            utcHipCheck(hipMalloc((void**)&m_dev, m_bytes), "GpuData malloc");
            m_status = MemStatus::host;
            break;
        case MemType::pinned:
            utcHipCheck(hipHostMalloc((void**)&m_host, m_bytes, hipHostMallocDefault), "GpuData host malloc");
            utcHipCheck(hipMalloc((void**)&m_dev, m_bytes), "GpuData malloc");
            m_status = MemStatus::host;
            break;
        case MemType::unified:
            utcHipCheck(hipMallocManaged((void**)&m_dev, m_bytes, hipMemAttachGlobal), "GpuData managed");
            m_host = m_dev;
            m_status = MemStatus::synced;
            break;
        default:
            std::fprintf(stderr, "ERROR, undefined gpu memory type\n");
            return;
        }
        m_init = true;
    }

    unsigned long m_size[3] = {0, 0, 0};
    unsigned long m_bytes = 0;
    int m_dim = 0;
    MemType m_memtype = MemType::pageable;
    MemStatus m_status = MemStatus::unknown;
    T* m_host = nullptr;
    T* m_dev = nullptr;
    bool m_init = false;
};

}  // namespace iUtc
