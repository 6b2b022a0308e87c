// runtime.hip -- device context, arena allocator, status, timers, copies and
// fills of libophip.
//
// Replaces the reference's runtime layer (oplib/opcuda/util.cu:59-290,
// sset.cu:41-74, scopy2.cu:83-144).  Differences, by design:
//   * device buffers are 32-bit offsets into one per-device arena (see
//     ophip_internal.h), not truncated 64-bit pointers;
//   * all work goes to one non-blocking stream per device so it can be
//     captured into a HIP graph; opcuda_thread_synchronize waits for it;
//   * the status word is thread-local.
#include "ophip_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <map>
#include <mutex>

namespace ophip {

namespace {

constexpr int kMaxDev = 16;
constexpr size_t kAlign = 256;
constexpr size_t kDefaultArena = (size_t)4 << 30;  // 4 GiB: 32-bit handles
constexpr size_t kMaxArena = ((size_t)4 << 30) - kAlign;

thread_local unsigned t_status = 0;
thread_local int t_dev = -1;

struct Arena {
    // free blocks: offset -> size; allocated blocks: offset -> size
    std::map<size_t, size_t> free_blocks;
    std::map<size_t, size_t> used;
};

Device g_dev[kMaxDev];
Arena g_arena[kMaxDev];
std::mutex g_mu;

size_t arena_size_from_env() {
    const char *s = getenv("OPHIP_ARENA_MB");
    if (!s) return kDefaultArena - kAlign;
    size_t mb = strtoull(s, nullptr, 10);
    size_t b = mb << 20;
    if (b == 0 || b > kMaxArena) b = kMaxArena;
    return b;
}

bool init_device(int id) {
    Device &d = g_dev[id];
    if (d.base) return true;
    hipError_t e = hipSetDevice(id);
    if (e != hipSuccess) { set_status(hip_status(e)); return false; }
    size_t bytes = arena_size_from_env();
    void *p = nullptr;
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) { set_status(hip_status(e)); return false; }
    e = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking);
    if (e != hipSuccess) { (void)hipFree(p); set_status(hip_status(e)); return false; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, id) == hipSuccess) d.n_cu = prop.multiProcessorCount;
    d.base = (char *)p;
    d.arena_bytes = bytes;
    d.id = id;
    // offset 0 is the null handle: the first usable byte is kAlign
    g_arena[id].free_blocks[kAlign] = bytes - kAlign;
    return true;
}

}  // namespace

unsigned set_status(unsigned s) {
    t_status = s;
    return s;
}

unsigned hip_status(hipError_t e) { return (unsigned)e; }

Device &dev() {
    if (t_dev < 0) {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) cur = 0;
        t_dev = cur;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    init_device(t_dev);
    return g_dev[t_dev];
}

Device *dev_ready() {
    int id = t_dev;
    if (id < 0 && hipGetDevice(&id) != hipSuccess) return nullptr;
    if (id < 0 || id >= kMaxDev) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    Device &d = g_dev[id];
    return d.base ? &d : nullptr;
}

void *ptr(uint32_t h) {
    if (h == 0) return nullptr;
    Device &d = dev();
    if (!d.base || (size_t)h >= d.arena_bytes) return nullptr;
    return d.base + h;
}

unsigned check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_status(hip_status(e));
    return set_status(kOk);
}

void *workspace(size_t bytes) {
    Device &d = dev();
    if (d.ws_bytes >= bytes) return d.ws;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (d.live_graphs > 0 || (hipStreamIsCapturing(d.stream, &cs) == hipSuccess &&
                              cs != hipStreamCaptureStatusNone)) {
        set_status(kErrUnsupported);  // a graph holds the current address
        return nullptr;
    }
    if (d.ws) {
        (void)hipStreamSynchronize(d.stream);
        (void)hipFree(d.ws);
        d.ws = nullptr;
        d.ws_bytes = 0;
    }
    size_t b = bytes < ((size_t)16 << 20) ? ((size_t)16 << 20) : bytes;
    if (hipMalloc(&d.ws, b) != hipSuccess) {
        set_status(kErrOutOfArena);
        return nullptr;
    }
    d.ws_bytes = b;
    return d.ws;
}

// arena allocation: first fit, 256-byte granules, coalescing free
static uint32_t arena_alloc(size_t sz) {
    Device &d = dev();
    if (!d.base) return 0;
    size_t need = (sz + kAlign - 1) & ~(kAlign - 1);
    if (need == 0) need = kAlign;
    std::lock_guard<std::mutex> lk(g_mu);
    Arena &a = g_arena[d.id];
    for (auto it = a.free_blocks.begin(); it != a.free_blocks.end(); ++it) {
        if (it->second >= need) {
            size_t off = it->first, rest = it->second - need;
            a.free_blocks.erase(it);
            if (rest) a.free_blocks[off + need] = rest;
            a.used[off] = need;
            return (uint32_t)off;
        }
    }
    return 0;
}

static bool arena_free(uint32_t h) {
    Device &d = dev();
    std::lock_guard<std::mutex> lk(g_mu);
    Arena &a = g_arena[d.id];
    auto it = a.used.find(h);
    if (it == a.used.end()) return false;
    size_t off = it->first, sz = it->second;
    a.used.erase(it);
    auto nx = a.free_blocks.lower_bound(off);
    if (nx != a.free_blocks.end() && off + sz == nx->first) {
        sz += nx->second;
        nx = a.free_blocks.erase(nx);
    }
    if (nx != a.free_blocks.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += sz;
            return true;
        }
    }
    a.free_blocks[off] = sz;
    return true;
}

// byte range [h, h+sz) must lie inside one allocated block
bool handle_range_ok(uint32_t h, size_t sz) {
    Device &d = dev();
    if (h == 0 || !d.base) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    Arena &a = g_arena[d.id];
    auto it = a.used.upper_bound(h);
    if (it == a.used.begin()) return false;
    --it;
    return (size_t)h + sz <= it->first + it->second;
}

}  // namespace ophip

using namespace ophip;

// ---------------------------------------------------------------------
// context, device, status (util.cu:115-160, 204)
// ---------------------------------------------------------------------

OPHIP_EXPORT int opcuda_cublas_init(void) {
    Device &d = dev();
    return (int)set_status(d.base ? kOk : kErrNoDevice);
}

OPHIP_EXPORT int opcuda_shutdown(void) {
    Device &d = dev();
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    return (int)set_status(kOk);
}

OPHIP_EXPORT void opcuda_thread_synchronize(void) {
    Device &d = dev();
    hipError_t e = hipStreamSynchronize(d.stream);
    if (e != hipSuccess) set_status(hip_status(e));
}

OPHIP_EXPORT void opcuda_set_device(int id) {
    if (id < 0 || id >= 16) { set_status(kErrBadArg); return; }
    hipError_t e = hipSetDevice(id);
    if (e != hipSuccess) { set_status(hip_status(e)); return; }
    t_dev = id;
}

OPHIP_EXPORT int opcuda_get_device(void) {
    dev();
    return t_dev;
}

OPHIP_EXPORT int opcuda_device_get_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

OPHIP_EXPORT int opcuda_device_count(void) { return opcuda_device_get_count(); }

OPHIP_EXPORT int opcuda_multi_processor_count(int id) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, id) != hipSuccess) return 0;
    return p.multiProcessorCount;
}

OPHIP_EXPORT int opcuda_max_threads_per_block(int id) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, id) != hipSuccess) return 0;
    return p.maxThreadsPerBlock;
}

OPHIP_EXPORT unsigned opcuda_get_status(void) { return t_status; }

// sets the calling thread's status (a layer above the library reports its
// own failures through it, e.g. after cleanup calls that succeeded)
OPHIP_EXPORT unsigned ophip_set_status(unsigned s) { return set_status(s); }

// ---------------------------------------------------------------------
// memory (util.cu:182-236)
// ---------------------------------------------------------------------

OPHIP_EXPORT unsigned opcuda_mem_alloc(unsigned sz) {
    Device &d = dev();
    if (!d.base) { set_status(kErrNoDevice); return 0; }
    uint32_t h = arena_alloc(sz);
    set_status(h ? kOk : kErrOutOfArena);
    return h;
}

OPHIP_EXPORT void opcuda_mem_free_device(unsigned h) {
    Device &d = dev();
    if (d.stream) (void)hipStreamSynchronize(d.stream);  // util.cu:219 syncs first
    set_status(arena_free(h) ? kOk : kErrBadHandle);
}

OPHIP_EXPORT unsigned long opcuda_mem_alloc_host(unsigned sz) {
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, sz, hipHostMallocDefault);
    set_status(hip_status(e));
    return (unsigned long)(uintptr_t)p;
}

OPHIP_EXPORT void opcuda_mem_free_host(unsigned long h) {
    set_status(hip_status(hipHostFree((void *)(uintptr_t)h)));
}

OPHIP_EXPORT int opcuda_memcpy_h2d(unsigned dh, unsigned long hp, unsigned sz) {
    Device &d = dev();
    if (!handle_range_ok(dh, sz)) return (int)set_status(kErrBadHandle);
    hipError_t e = hipMemcpyAsync(ptr(dh), (const void *)(uintptr_t)hp, sz,
                                  hipMemcpyHostToDevice, d.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d.stream);
    return (int)set_status(hip_status(e));
}

OPHIP_EXPORT int opcuda_memcpy_d2h(unsigned dh, unsigned long hp, unsigned sz) {
    Device &d = dev();
    if (!handle_range_ok(dh, sz)) return (int)set_status(kErrBadHandle);
    hipError_t e = hipMemcpyAsync((void *)(uintptr_t)hp, ptr(dh), sz,
                                  hipMemcpyDeviceToHost, d.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d.stream);
    return (int)set_status(hip_status(e));
}

OPHIP_EXPORT int opcuda_memcpy_d2d(unsigned dst, unsigned src, unsigned sz) {
    Device &d = dev();
    if (!handle_range_ok(dst, sz) || !handle_range_ok(src, sz))
        return (int)set_status(kErrBadHandle);
    hipError_t e = hipMemcpyAsync(ptr(dst), ptr(src), sz, hipMemcpyDeviceToDevice, d.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d.stream);
    return (int)set_status(hip_status(e));
}

// ---------------------------------------------------------------------
// timers (util.cu:62-113): wall clock, as the reference's Linux build
// ---------------------------------------------------------------------

static double g_t0 = 0.0, g_t1 = 0.0;
static double now_s() {
    timeval tp;
    gettimeofday(&tp, nullptr);
    return (double)tp.tv_sec + (double)tp.tv_usec * 1e-6;
}
OPHIP_EXPORT void opcuda_create_timer(void) { g_t0 = now_s(); }
OPHIP_EXPORT void opcuda_reset_timer(void) { g_t0 = now_s(); }
OPHIP_EXPORT void opcuda_stop_timer(void) { g_t1 = now_s(); }
OPHIP_EXPORT double opcuda_get_timer_value(void) { return g_t1 - g_t0; }

// ---------------------------------------------------------------------
// strided copy and fills (util.cu:265-283, sset.cu:41-74, scopy2.cu:83-144)
// ---------------------------------------------------------------------

__global__ void k_scopy(int n, const float *__restrict__ x, int incx, float *__restrict__ y,
                        int incy) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        y[(size_t)j * incy] = x[(size_t)j * incx];
}

__global__ void k_copy4(size_t n4, const float4 *__restrict__ x, float4 *__restrict__ y) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < n4;
         j += (size_t)gridDim.x * blockDim.x)
        y[j] = x[j];
}

__global__ void k_copy1(size_t n, size_t from, const float *__restrict__ x, float *__restrict__ y) {
    size_t j = from + blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (j < n) y[j] = x[j];
}

// read-bandwidth probe: every 16-byte word of x read once, four loads in
// flight per thread, folded by XOR; the store depends on the data (a
// pattern random inputs essentially never give), so no load is dead and
// almost nothing is written
constexpr int kProbeGrid = 2048;
__global__ __launch_bounds__(256) void k_read16(size_t n16, const uint4 *__restrict__ x,
                                                uint4 *__restrict__ out) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    uint4 a = {0u, 0u, 0u, 0u};
    for (; j + 3 * stride < n16; j += 4 * stride) {
        const uint4 v0 = x[j], v1 = x[j + stride], v2 = x[j + 2 * stride], v3 = x[j + 3 * stride];
        a.x ^= v0.x ^ v1.x ^ v2.x ^ v3.x;
        a.y ^= v0.y ^ v1.y ^ v2.y ^ v3.y;
        a.z ^= v0.z ^ v1.z ^ v2.z ^ v3.z;
        a.w ^= v0.w ^ v1.w ^ v2.w ^ v3.w;
    }
    for (; j < n16; j += stride) {
        const uint4 v = x[j];
        a.x ^= v.x;
        a.y ^= v.y;
        a.z ^= v.z;
        a.w ^= v.w;
    }
    if ((a.x ^ a.y ^ a.z ^ a.w) == 0x9e3779b9u) out[blockIdx.x] = a;
}

__global__ void k_sset(int n, float c, float *__restrict__ x, int incx) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        x[(size_t)j * incx] = c;
}

static int grid_for(size_t n, int bs) {
    size_t g = (n + bs - 1) / bs;
    if (g > 65536) g = 65536;
    if (g == 0) g = 1;
    return (int)g;
}

OPHIP_EXPORT void opcuda_scopy(int n, unsigned x, int incx, unsigned y, int incy) {
    Device &d = dev();
    if (n <= 0) { set_status(kOk); return; }
    if (!ptr(x) || !ptr(y)) { set_status(kErrBadHandle); return; }
    hipLaunchKernelGGL(k_scopy, dim3(grid_for(n, 256)), dim3(256), 0, d.stream, n,
                       P<const float>(x), incx, P<float>(y), incy);
    check_launch();
}

// Contiguous copy of n floats.  The reference returns void (scopy2.cu:137)
// while its C# caller declares int and throws on non-zero (CModel
// Device.cs:74, 289-292); this returns the status, 0 on success.  Unlike
// global_scopy1 it never writes past n (scopy2.cu:91 tests blockDim for
// the last block).
OPHIP_EXPORT int opcuda_scopy1(unsigned dst, unsigned src, unsigned n) {
    Device &d = dev();
    if (n == 0) return (int)set_status(kOk);
    if (!handle_range_ok(dst, (size_t)n * 4) || !handle_range_ok(src, (size_t)n * 4))
        return (int)set_status(kErrBadHandle);
    const float *xs = P<const float>(src);
    float *ys = P<float>(dst);
    size_t n4 = ((((uintptr_t)xs | (uintptr_t)ys) & 15) == 0) ? n / 4 : 0;
    if (n4) {
        hipLaunchKernelGGL(k_copy4, dim3(grid_for(n4, 256)), dim3(256), 0, d.stream, n4,
                           (const float4 *)xs, (float4 *)ys);
    }
    size_t rest = n - n4 * 4;
    if (rest) {
        hipLaunchKernelGGL(k_copy1, dim3(grid_for(rest, 256)), dim3(256), 0, d.stream,
                           (size_t)n, n4 * 4, xs, ys);
    }
    return (int)check_launch();
}

// Read-bandwidth probe (measurement; no reference counterpart): reads the
// bytes (a multiple of 16) at handle x once on the library stream; out:
// 32 KiB that a data-dependent store may touch.  Timed between events, a
// buffer that fits the Infinity Cache and was read just before gives the
// cache's read bandwidth, a larger one HBM's.
OPHIP_EXPORT int ophip_read_probe(unsigned x, unsigned long bytes, unsigned out) {
    Device &d = dev();
    if (bytes == 0 || bytes % 16) return (int)set_status(kErrBadArg);
    if (!handle_range_ok(x, bytes) || !handle_range_ok(out, (size_t)16 * kProbeGrid))
        return (int)set_status(kErrBadHandle);
    hipLaunchKernelGGL(k_read16, dim3(kProbeGrid), dim3(256), 0, d.stream, (size_t)bytes / 16,
                       P<const uint4>(x), P<uint4>(out));
    return (int)check_launch();
}

OPHIP_EXPORT void opcuda_ssetall(unsigned x, int n, float c, int incx) {
    Device &d = dev();
    if (n <= 0) { set_status(kOk); return; }
    if (!ptr(x)) { set_status(kErrBadHandle); return; }
    hipLaunchKernelGGL(k_sset, dim3(grid_for(n, 256)), dim3(256), 0, d.stream, n, c,
                       P<float>(x), incx);
    check_launch();
}

// x[i] = c (sset.cu:61-72).  `n` is the buffer length and is checked.
OPHIP_EXPORT void opcuda_ssetone(unsigned x, int n, float c, int i) {
    Device &d = dev();
    if (i < 0 || i >= n) { set_status(kErrBadArg); return; }
    if (!ptr(x)) { set_status(kErrBadHandle); return; }
    hipLaunchKernelGGL(k_sset, dim3(1), dim3(64), 0, d.stream, 1, c, P<float>(x) + i, 1);
    check_launch();
}

__global__ void k_axpby(size_t n, float a, const float *__restrict__ A, float b,
                        const float *__restrict__ B, float *__restrict__ C) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < n;
         j += (size_t)gridDim.x * blockDim.x)
        C[j] = a * A[j] + b * B[j];
}

OPHIP_EXPORT int ophip_saxpby(int n, float a, unsigned A, float b, unsigned B, unsigned C) {
    Device &d = dev();
    if (n < 0) return (int)set_status(kErrBadArg);
    if (n == 0) return (int)set_status(kOk);
    if (!handle_range_ok(A, (size_t)n * 4) || !handle_range_ok(B, (size_t)n * 4) ||
        !handle_range_ok(C, (size_t)n * 4))
        return (int)set_status(kErrBadHandle);
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256)), dim3(256), 0, d.stream, (size_t)n, a,
                       P<const float>(A), b, P<const float>(B), P<float>(C));
    return (int)check_launch();
}

// ---------------------------------------------------------------------
// extensions: raw device address of a handle (for zero-copy interop with
// torch tensors allocated elsewhere) and arena statistics
// ---------------------------------------------------------------------

OPHIP_EXPORT unsigned long ophip_device_address(unsigned h) {
    return (unsigned long)(uintptr_t)ptr(h);
}

OPHIP_EXPORT unsigned long ophip_arena_bytes(void) { return dev().arena_bytes; }

OPHIP_EXPORT unsigned long ophip_stream(void) { return (unsigned long)(uintptr_t)dev().stream; }
