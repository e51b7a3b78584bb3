// Internal definitions behind the opaque C handles, shared by capi.cpp and sched.cpp.
#pragma once
#include "engine.h"
#include <atomic>
#include <functional>

struct gsj_ctx { gsj::Context c; };

// A layer is reference counted: the owner's gsj_layer_free drops one reference, and every request queued or running
// on the scheduler holds one, so device memory outlives the requests that read it.
struct gsj_layer {
  gsj::Layer L;
  std::atomic<int> refs{1};
};

namespace gsj {
extern thread_local std::string g_last_error;
// run f, mapping exceptions to GSJ_E_* codes and the thread's last-error message
int guarded_status(const std::function<void()>& f);
// device, stream (non-blocking, default priority) and timing events of a context
void context_init(Context& c, int device);
// drop one reference; the last one frees the layer (after its owner context's stream drains)
void layer_release(gsj_layer* layer);
// once per device, with it current: the address of the no-allocation entry point's scratch on it (noalloc.hip)
void noalloc_setup(int device);
}  // namespace gsj
