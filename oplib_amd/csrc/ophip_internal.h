// ophip_internal.h -- shared plumbing of libophip (not part of the C-ABI).
//
// Device memory model: one arena per device, reserved with a single
// hipMalloc.  Every device buffer crosses the C-ABI as a 32-bit byte offset
// into that arena ("handle"); 0 is the null handle.  Kernels receive the
// arena base and add it to every handle, including handles stored inside
// device-resident pointer tables and block descriptors.  This keeps the
// reference's 32-bit "device pointer" ABI (CArray.cs:46, util.cu:209-214)
// and its caller-side pointer arithmetic (CModel Device.cs:193) valid while
// addressing up to 4 GiB of HBM per device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace ophip {

// status codes beyond hipError_t (kept above hipErrorUnknown's range)
enum : unsigned {
    kOk = 0,
    kErrBadHandle = 0x10001,
    kErrOutOfArena = 0x10002,
    kErrBadArg = 0x10003,
    kErrNoDevice = 0x10004,
    kErrUnsupported = 0x10005,
};

struct Device {
    char *base = nullptr;      // arena base
    size_t arena_bytes = 0;    // reserved size
    hipStream_t stream = nullptr;
    int id = -1;
    int n_cu = 0;
    void *ws = nullptr;        // library workspace (split-K partials etc.)
    size_t ws_bytes = 0;
    int live_graphs = 0;       // instantiated graphs: they hold the workspace address
};

// current device context (creates the arena and stream on first use)
Device &dev();
// the current device context if its arena and stream exist, else nullptr;
// never creates them (no allocation)
Device *dev_ready();
// handle -> device address; nullptr for 0 or out-of-arena handles
void *ptr(uint32_t h);
template <class T> inline T *P(uint32_t h) { return reinterpret_cast<T *>(ptr(h)); }
// set the thread-local status and return it
unsigned set_status(unsigned s);
unsigned hip_status(hipError_t e);
// check the last launch, record its status
unsigned check_launch();
// true when [h, h+bytes) lies inside one live arena allocation
bool handle_range_ok(uint32_t h, size_t bytes);
// library-owned device workspace of at least `bytes` (grown on demand,
// never shrunk; only call between kernels on the library stream).  It is
// not grown while a graph is being captured or any instantiated graph is
// alive (those hold its address): reserve enough before capturing.
void *workspace(size_t bytes);

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace ophip

#define OPHIP_EXPORT extern "C" __attribute__((visibility("default")))
