// graph.hip -- HIP-graph capture of the library stream, stream events for
// timing, and workspace reservation.  No reference counterpart: the
// reference launches on the legacy default stream (CModel Device.cs,
// util.cu:131) and times with a host wall clock (util.cu:62-113).
#include "ophip_internal.h"

using namespace ophip;

namespace {
constexpr int kSlots = 64;
hipEvent_t g_ev[kSlots] = {};
}  // namespace

// Grow the library workspace now so that a later capture never allocates.
OPHIP_EXPORT int ophip_reserve_workspace(unsigned long bytes) {
    return workspace((size_t)bytes) ? (int)set_status(kOk) : (int)kErrOutOfArena;
}

// Start capturing every launch on the library stream into a graph.
OPHIP_EXPORT int ophip_capture_begin(void) {
    Device &D = dev();
    return (int)set_status(hip_status(hipStreamBeginCapture(D.stream, hipStreamCaptureModeThreadLocal)));
}

// End the capture and instantiate; returns an opaque executable-graph
// handle (0 on failure, status set).
OPHIP_EXPORT unsigned long ophip_capture_end(void) {
    Device &D = dev();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(D.stream, &g);
    if (e != hipSuccess || !g) { set_status(hip_status(e)); return 0; }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { set_status(hip_status(e)); return 0; }
    D.live_graphs++;
    set_status(kOk);
    return (unsigned long)(uintptr_t)ex;
}

OPHIP_EXPORT int ophip_graph_launch(unsigned long gx) {
    Device &D = dev();
    if (!gx) return (int)set_status(kErrBadArg);
    return (int)set_status(hip_status(hipGraphLaunch((hipGraphExec_t)(uintptr_t)gx, D.stream)));
}

OPHIP_EXPORT void ophip_graph_destroy(unsigned long gx) {
    if (!gx) return;
    Device &D = dev();
    (void)hipStreamSynchronize(D.stream);
    (void)hipGraphExecDestroy((hipGraphExec_t)(uintptr_t)gx);
    if (D.live_graphs > 0) D.live_graphs--;
}

// Record event `slot` on the library stream (the stream every kernel of
// this library is launched on).
OPHIP_EXPORT int ophip_event_record(int slot) {
    Device &D = dev();
    if (slot < 0 || slot >= kSlots) return (int)set_status(kErrBadArg);
    if (!g_ev[slot]) {
        hipError_t e = hipEventCreate(&g_ev[slot]);
        if (e != hipSuccess) return (int)set_status(hip_status(e));
    }
    return (int)set_status(hip_status(hipEventRecord(g_ev[slot], D.stream)));
}

// Milliseconds between two recorded events (waits for the second one).
OPHIP_EXPORT double ophip_event_elapsed_ms(int a, int b) {
    if (a < 0 || b < 0 || a >= kSlots || b >= kSlots || !g_ev[a] || !g_ev[b]) {
        set_status(kErrBadArg);
        return -1.0;
    }
    (void)hipEventSynchronize(g_ev[b]);
    float ms = 0.0f;
    hipError_t e = hipEventElapsedTime(&ms, g_ev[a], g_ev[b]);
    set_status(hip_status(e));
    return e == hipSuccess ? (double)ms : -1.0;
}
