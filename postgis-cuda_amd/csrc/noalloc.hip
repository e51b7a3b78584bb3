// gsj_join_noalloc: the join / relate entry point without any allocation, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid it). Every scratch buffer of the join path comes from
// one fixed device array per GPU (`scratch` below, NOALLOC_SCRATCH bytes, held by each GPU the library runs on until the
// process ends) and one fixed host region per GPU (set up with the context); the kernels are the join's own, so the rows
// are bit-identical to gsj_join's.
//
// The array is shared by every call on its GPU and keeps the last call's contents: the join path clears what must start
// at zero (counters, queue controls) itself, as it does for its grow-only buffers. Calls on one GPU are serialised by a
// per-GPU lock; overlapping calls would race on the array.
#include "capi_internal.h"
#include <memory>

namespace {

constexpr size_t NOALLOC_SCRATCH = 128ull << 20;      // device bytes per GPU
constexpr size_t NOALLOC_HOST = 16ull << 20;          // host bytes per GPU (row staging, scalars)
constexpr uint64_t NOALLOC_MAX_PAIRS = 128ull << 10;  // candidate pairs per sub-batch (default and maximum)
constexpr uint32_t NOALLOC_MAX_PROBES = 32u << 10;    // probes per traversal batch (default and maximum)
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH];

struct PerDevice {
  std::once_flag once;
  unsigned char* dev = nullptr;            // this GPU's copy of `scratch` (hipGetSymbolAddress with the GPU current)
  std::unique_ptr<unsigned char[]> host;   // host staging region
  int setup_rc = GSJ_OK;
  std::string setup_msg;
  std::mutex mu;                           // one call at a time per GPU: they share the array
};
PerDevice g_dev[NOALLOC_MAX_DEVICES];

}  // namespace

namespace gsj {

void noalloc_setup(int device) {
  GSJ_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, GSJ_E_LIMIT, "device index beyond the no-allocation table");
  PerDevice& d = g_dev[device];
  std::call_once(d.once, [&] {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));   // the caller made `device` current
    if (e != hipSuccess || !p) {
      d.setup_rc = GSJ_E_HIP;
      d.setup_msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return;
    }
    d.dev = static_cast<unsigned char*>(p);
    d.host.reset(new unsigned char[NOALLOC_HOST]);
  });
}

}  // namespace gsj

extern "C" int gsj_join_noalloc(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, const gsj_join_opts* opts, gsj_rows_cb cb, void* user,
                                gsj_join_stats* stats) {
  return gsj::guarded_status([&] {
    GSJ_REQUIRE(ctx && P && R, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(P->L.ctx == &ctx->c && R->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    gsj_join_opts o{};
    if (opts) o = *opts;
    GSJ_REQUIRE(!o.has_window, GSJ_E_ARG, "gsj_join_noalloc: query windows need per-layer scratch; use gsj_join");
    GSJ_REQUIRE(o.max_pairs <= NOALLOC_MAX_PAIRS, GSJ_E_LIMIT,
                "gsj_join_noalloc: max_pairs " + std::to_string(o.max_pairs) + " exceeds " + std::to_string(NOALLOC_MAX_PAIRS));
    GSJ_REQUIRE(o.batch_q <= NOALLOC_MAX_PROBES, GSJ_E_LIMIT,
                "gsj_join_noalloc: batch_q " + std::to_string(o.batch_q) + " exceeds " + std::to_string(NOALLOC_MAX_PROBES));
    if (!o.max_pairs) o.max_pairs = NOALLOC_MAX_PAIRS;
    if (!o.batch_q) o.batch_q = NOALLOC_MAX_PROBES;
    gsj::Context& owner = ctx->c;
    GSJ_REQUIRE(owner.device >= 0 && owner.device < NOALLOC_MAX_DEVICES, GSJ_E_LIMIT, "device index beyond the no-allocation table");
    PerDevice& d = g_dev[owner.device];
    GSJ_REQUIRE(d.setup_rc == GSJ_OK && d.dev, d.setup_rc == GSJ_OK ? GSJ_E_INTERNAL : d.setup_rc,
                d.setup_msg.empty() ? std::string("no-allocation scratch not set up") : d.setup_msg);
    std::lock_guard<std::mutex> lk(owner.mu);
    std::lock_guard<std::mutex> lk_dev(d.mu);
    GSJ_HIP(hipSetDevice(owner.device));
    gsj::Arena dev{d.dev, NOALLOC_SCRATCH, 0, "device scratch"}, host{d.host.get(), NOALLOC_HOST, 0, "host staging"};
    // a context of the caller's stream and events whose every scratch buffer is a slice of the arenas
    gsj::Context c;
    c.device = owner.device;
    c.stream = owner.stream;
    c.side = owner.side;
    for (size_t i = 0; i < sizeof c.ev / sizeof c.ev[0]; i++) c.ev[i] = owner.ev[i];
    c.cu_count = owner.cu_count;
    c.tune_from = &owner.tune;
    c.borrowed = true;
    c.dev_arena = &dev;
    c.for_each_devbuf([&](gsj::DevBuf& b) { b.arena = &dev; });
    c.for_each_pinbuf([&](gsj::PinBuf& b) { b.arena = &host; });
    gsj::join_run(c, P->L, R->L, o, cb, user, stats, nullptr);
  });
}
