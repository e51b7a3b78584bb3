// extern "C" boundary of libgsj (declared in include/gsj.h). Every entry point converts exceptions into GSJ_E_*
// codes plus a thread-local message; nothing here exits the process.
#include "capi_internal.h"
#include "hostprep.h"
#include <cstdlib>
#include <cstring>
#include <new>
#include <sstream>

struct gsj_hostset { gsj::HostSet s; };

namespace gsj {

thread_local std::string g_last_error;

int guarded_status(const std::function<void()>& f) {
  try {
    f();
    g_last_error.clear();
    return GSJ_OK;
  } catch (const gsj::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return GSJ_E_NOMEM;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return GSJ_E_ARG;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GSJ_E_IO;
  } catch (...) {
    g_last_error = "unknown error";
    return GSJ_E_INTERNAL;
  }
}

Context::~Context() {
  if (borrowed) return;
  for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
}

void context_init(Context& c, int device) {
  int ndev = 0;
  GSJ_HIP(hipGetDeviceCount(&ndev));
  GSJ_REQUIRE(device >= 0 && device < ndev, GSJ_E_ARG, "no such HIP device: " + std::to_string(device));
  GSJ_HIP(hipSetDevice(device));
  c.device = device;
  hipDeviceProp_t p;
  std::string arch;
  if (hipGetDeviceProperties(&p, device) == hipSuccess) {
    c.cu_count = p.multiProcessorCount;
    arch = p.gcnArchName;
    arch = arch.substr(0, arch.find(':'));   // "gfx950:sramecc+:xnack-" -> "gfx950"
  }
  c.tune = load_tuning(arch);
  GSJ_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  GSJ_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
  for (hipEvent_t& e : c.ev) GSJ_HIP(hipEventCreate(&e));
  noalloc_setup(device);
}

void layer_release(gsj_layer* layer) {
  if (!layer || layer->refs.fetch_sub(1) != 1) return;
  if (layer->L.ctx) {
    std::lock_guard<std::mutex> lk(layer->L.ctx->mu);
    (void)hipSetDevice(layer->L.ctx->device);
    (void)hipStreamSynchronize(layer->L.ctx->stream);
    delete layer;
    return;
  }
  delete layer;
}

}  // namespace gsj

namespace {
template <class F>
int guarded(F&& f) { return gsj::guarded_status(std::function<void()>(std::forward<F>(f))); }
void set_device(const gsj::Context& c) { GSJ_HIP(hipSetDevice(c.device)); }
using gsj::g_last_error;
}  // namespace

extern "C" {

const char* gsj_last_error(void) { return g_last_error.c_str(); }
const char* gsj_version(void) { return "gsj 0.1 (gfx950)"; }

int gsj_ctx_create(int device, gsj_ctx** out) {
  return guarded([&] {
    GSJ_REQUIRE(out, GSJ_E_ARG, "null output pointer");
    *out = nullptr;
    auto ctx = std::make_unique<gsj_ctx>();
    gsj::context_init(ctx->c, device);
    *out = ctx.release();
  });
}

void gsj_ctx_destroy(gsj_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  delete ctx;
}

const char* gsj_ctx_status_json(gsj_ctx* ctx) {
  if (!ctx) return "{}";
  std::lock_guard<std::mutex> lk(ctx->c.mu);
  int rc = guarded([&] {
    set_device(ctx->c);
    hipDeviceProp_t p;
    GSJ_HIP(hipGetDeviceProperties(&p, ctx->c.device));
    size_t fr = 0, tot = 0;
    GSJ_HIP(hipMemGetInfo(&fr, &tot));
    std::ostringstream o;
    o << "{\"device\":\"" << p.name << "\",\"arch\":\"" << p.gcnArchName << "\",\"cu\":" << p.multiProcessorCount
      << ",\"mem_total_mb\":" << (double)tot / 1048576.0 << ",\"mem_free_mb\":" << (double)fr / 1048576.0 << "}";
    ctx->c.status = o.str();
  });
  if (rc != GSJ_OK) ctx->c.status = "{}";
  return ctx->c.status.c_str();
}

int gsj_ctx_set_tuning(gsj_ctx* ctx, const char* text) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && text, GSJ_E_ARG, "null context or text");
    gsj::Tuning t = gsj::Tuning::parse(text, "gsj_ctx_set_tuning");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    ctx->c.tune = std::move(t);
  });
}

const char* gsj_ctx_tuning(gsj_ctx* ctx) {
  if (!ctx) return "";
  std::lock_guard<std::mutex> lk(ctx->c.mu);
  ctx->c.tune_text = ctx->c.tune.text();
  return ctx->c.tune_text.c_str();
}

int gsj_ctx_trim(gsj_ctx* ctx) {
  return guarded([&] {
    GSJ_REQUIRE(ctx, GSJ_E_ARG, "null context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    GSJ_HIP(hipStreamSynchronize(ctx->c.stream));
    ctx->c.for_each_devbuf([](gsj::DevBuf& b) { b.release(); });
  });
}

int gsj_layer_create_polys(gsj_ctx* ctx, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv,
                           const uint32_t* parent, const uint8_t* syn, gsj_layer** out) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && out, GSJ_E_ARG, "null context or output pointer");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    auto* l = new gsj_layer();
    l->L.ctx = &ctx->c;
    try {
      gsj::layer_upload_polys(l->L, gro, rvo, xy, n, nr, nv, parent, syn);
    } catch (...) {
      delete l;
      throw;
    }
    *out = l;
  });
}

int gsj_layer_create_points(gsj_ctx* ctx, const double* xy, uint64_t n, gsj_layer** out) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && out, GSJ_E_ARG, "null context or output pointer");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    auto* l = new gsj_layer();
    l->L.ctx = &ctx->c;
    try {
      gsj::layer_upload_points(l->L, xy, n);
    } catch (...) {
      delete l;
      throw;
    }
    *out = l;
  });
}

void gsj_layer_free(gsj_layer* layer) { gsj::layer_release(layer); }

int gsj_bvh_build(gsj_layer* layer) {
  return guarded([&] {
    GSJ_REQUIRE(layer, GSJ_E_ARG, "null layer");
    std::lock_guard<std::mutex> lk(layer->L.ctx->mu);
    set_device(*layer->L.ctx);
    gsj::bvh_build(layer->L);
  });
}

int gsj_layer_info_get(const gsj_layer* layer, gsj_layer_info* info) {
  return guarded([&] {
    GSJ_REQUIRE(layer && info, GSJ_E_ARG, "null layer or info");
    const gsj::Layer& L = layer->L;
    info->n = L.n; info->nr = L.nr; info->nv = L.nv; info->nparent = L.nparent;
    info->is_points = L.points ? 1 : 0; info->has_bvh = L.has_bvh ? 1 : 0;
    info->device_bytes = L.mem_bytes + L.bvh_bytes;
    info->bvh_ms = L.build_ms;
  });
}

int gsj_layer_export(const gsj_layer* layer, double* origin, double* ext, float* bb, double* q, double* blen) {
  return guarded([&] {
    GSJ_REQUIRE(layer, GSJ_E_ARG, "null layer");
    const gsj::Layer& L = layer->L;
    std::lock_guard<std::mutex> lk(L.ctx->mu);
    set_device(*L.ctx);
    hipStream_t s = L.ctx->stream;
    if (L.points) {
      if (bb && L.n) GSJ_HIP(hipMemcpyAsync(bb, L.pts.bb, L.n * 16, hipMemcpyDeviceToHost, s));
    } else {
      if (origin && L.n) GSJ_HIP(hipMemcpyAsync(origin, L.poly.origin, L.n * 16, hipMemcpyDeviceToHost, s));
      if (ext && L.n) GSJ_HIP(hipMemcpyAsync(ext, L.poly.ext, L.n * 16, hipMemcpyDeviceToHost, s));
      if (bb && L.n) GSJ_HIP(hipMemcpyAsync(bb, L.poly.bb, L.n * 16, hipMemcpyDeviceToHost, s));
      if (q && L.nv) GSJ_HIP(hipMemcpyAsync(q, L.poly.xy, L.nv * 16, hipMemcpyDeviceToHost, s));
      if (blen && L.nparent) GSJ_HIP(hipMemcpyAsync(blen, L.poly.blen, (size_t)L.nparent * 8, hipMemcpyDeviceToHost, s));
    }
    GSJ_HIP(hipStreamSynchronize(s));
  });
}

static gsj_join_opts defaults(const gsj_join_opts* o) {
  gsj_join_opts d{};
  if (o) d = *o;
  return d;
}

int gsj_join(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, const gsj_join_opts* opts, gsj_rows_cb cb, void* user, gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && P && R, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(P->L.ctx == &ctx->c && R->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::join_run(ctx->c, P->L, R->L, defaults(opts), cb, user, stats, nullptr);
  });
}

int gsj_points_in_polygons(gsj_ctx* ctx, const gsj_layer* pts, const gsj_layer* R, const gsj_join_opts* opts, gsj_points_cb cb, void* user,
                           gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && pts && R, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(pts->L.ctx == &ctx->c && R->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::points_join_run(ctx->c, pts->L, R->L, defaults(opts), cb, user, stats);
  });
}

int gsj_dwithin(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, double d, int same, int report_distance, const gsj_join_opts* opts,
                gsj_dist_cb cb, void* user, gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && P && R, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(P->L.ctx == &ctx->c && R->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::dwithin_run(ctx->c, P->L, R->L, d, same != 0, report_distance != 0, defaults(opts), cb, user, stats);
  });
}

int gsj_points_dwithin(gsj_ctx* ctx, const gsj_layer* pts, const gsj_layer* R, double d, int report_distance, const gsj_join_opts* opts,
                       gsj_dist_cb cb, void* user, gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && pts && R, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(pts->L.ctx == &ctx->c && R->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::points_dwithin_run(ctx->c, pts->L, R->L, d, report_distance != 0, defaults(opts), cb, user, stats);
  });
}

int gsj_intersects_buffered(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, double ra, double rb, int same, const uint8_t* need_self,
                            const gsj_join_opts* opts, gsj_dist_cb cb, void* user, gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && P && R, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(P->L.ctx == &ctx->c && R->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::SelfNeedHost need;
    need.flags = need_self;
    gsj::erode_run(ctx->c, P->L, R->L, ra, rb, same != 0, need, defaults(opts), cb, user, stats);
  });
}

int gsj_points_intersects_buffered(gsj_ctx* ctx, const gsj_layer* pts, const gsj_layer* R, double r, const gsj_join_opts* opts, gsj_dist_cb cb,
                                   void* user, gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && pts && R, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(pts->L.ctx == &ctx->c && R->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::points_erode_run(ctx->c, pts->L, R->L, r, defaults(opts), cb, user, stats);
  });
}

int gsj_cluster(gsj_ctx* ctx, const gsj_layer* P, const gsj_join_opts* opts, uint32_t* root, uint64_t nroot, gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && P, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(P->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::cluster_run(ctx->c, P->L, defaults(opts), root, nroot, stats);
  });
}

int gsj_cluster_rows(gsj_ctx* ctx, const gsj_layer* P, const gsj_join_opts* opts, gsj_rows_cb cb, void* user, gsj_join_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && P, GSJ_E_ARG, "null context or layer");
    GSJ_REQUIRE(P->L.ctx == &ctx->c, GSJ_E_ARG, "layer belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::cluster_rows_run(ctx->c, P->L, defaults(opts), cb, user, stats);
  });
}

void gsj_de9im_str(uint32_t code, char out[9]) {
  for (int k = 0; k < 9; k++) out[k] = "F012"[(code >> (2 * k)) & 3];
}
uint32_t gsj_de9im_transpose(uint32_t c) {
  uint32_t t = 0;
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++) t |= ((c >> (2 * (3 * r + q))) & 3u) << (2 * (3 * q + r));
  return t;
}
int gsj_de9im_match(uint32_t code, const char* pat) {
  if (!pat) return 0;
  for (int k = 0; k < 9; k++) {
    const char p = pat[k];
    const uint32_t v = (code >> (2 * k)) & 3u;
    if (p == '*') continue;
    if (p == 'T' || p == 't') { if (v == 0) return 0; continue; }
    if (p == 'F' || p == 'f') { if (v != 0) return 0; continue; }
    if (p >= '0' && p <= '2') { if (v != (uint32_t)(p - '0') + 1) return 0; continue; }
    return 0;
  }
  return 1;
}

// ---- host sets
void gsj_hostset_free(gsj_hostset* s) { delete s; }
int gsj_hostset_sizes(const gsj_hostset* s, uint64_t* n, uint64_t* nr, uint64_t* nv) {
  return guarded([&] {
    GSJ_REQUIRE(s, GSJ_E_ARG, "null set");
    if (n) *n = s->s.n();
    if (nr) *nr = s->s.nr();
    if (nv) *nv = s->s.nv();
  });
}
int gsj_hostset_arrays(const gsj_hostset* s, const uint64_t** gro, const uint64_t** rvo, const double** xy, const uint32_t** parent, const uint8_t** syn) {
  return guarded([&] {
    GSJ_REQUIRE(s, GSJ_E_ARG, "null set");
    if (gro) *gro = s->s.gro.data();
    if (rvo) *rvo = s->s.rvo.data();
    if (xy) *xy = s->s.xy.data();
    if (parent) *parent = s->s.has_parent ? s->s.parent.data() : nullptr;
    if (syn) *syn = s->s.has_parent ? s->s.syn.data() : nullptr;
  });
}

int gsj_gset_read(const char* path, gsj_hostset** out) {
  return guarded([&] {
    GSJ_REQUIRE(path && out, GSJ_E_ARG, "null path or output");
    try {
      auto* h = new gsj_hostset();
      try { h->s = gsj::read_gset(path); } catch (...) { delete h; throw; }
      *out = h;
    } catch (const std::runtime_error& e) {
      throw gsj::Error(GSJ_E_IO, e.what());
    }
  });
}
int gsj_gset_write(const char* path, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv) {
  return guarded([&] {
    GSJ_REQUIRE(path && gro && rvo, GSJ_E_ARG, "null argument");
    try { gsj::write_gset(path, gro, rvo, xy, n, nr, nv); } catch (const std::runtime_error& e) { throw gsj::Error(GSJ_E_IO, e.what()); }
  });
}
int gsj_pset_read(const char* path, double** xy, uint64_t* n) {
  return guarded([&] {
    GSJ_REQUIRE(path && xy && n, GSJ_E_ARG, "null argument");
    std::vector<double> v;
    try { v = gsj::read_pset(path); } catch (const std::runtime_error& e) { throw gsj::Error(GSJ_E_IO, e.what()); }
    *n = v.size() / 2;
    *xy = static_cast<double*>(std::malloc(std::max<size_t>(v.size(), 1) * 8));
    if (!*xy) throw std::bad_alloc();
    std::memcpy(*xy, v.data(), v.size() * 8);
  });
}
int gsj_pset_write(const char* path, const double* xy, uint64_t n) {
  return guarded([&] {
    GSJ_REQUIRE(path && (xy || n == 0), GSJ_E_ARG, "null argument");
    try { gsj::write_pset(path, xy, n); } catch (const std::runtime_error& e) { throw gsj::Error(GSJ_E_IO, e.what()); }
  });
}
void gsj_free(void* p) { std::free(p); }

int gsj_subdivide(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv, gsj_hostset** out) {
  return guarded([&] {
    GSJ_REQUIRE(gro && rvo && out, GSJ_E_ARG, "null argument");
    GSJ_REQUIRE(maxv >= 4, GSJ_E_ARG, "subdivide: maxv must be >= 4");
    gsj::check_offsets(gro, rvo, n);
    GSJ_REQUIRE(xy || rvo[gro[n]] == 0, GSJ_E_ARG, "null xy");
    auto* h = new gsj_hostset();
    h->s = gsj::subdivide(gro, rvo, xy, n, maxv);
    *out = h;
  });
}

int gsj_dedup(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, gsj_hostset** distinct, uint64_t* moff, uint32_t* midx,
              uint64_t* n_distinct) {
  return guarded([&] {
    GSJ_REQUIRE(gro && rvo && distinct && moff && midx && n_distinct, GSJ_E_ARG, "null argument");
    gsj::check_offsets(gro, rvo, n);
    GSJ_REQUIRE(xy || rvo[gro[n]] == 0, GSJ_E_ARG, "null xy");
    std::vector<uint64_t> mo;
    std::vector<uint32_t> mi;
    auto* h = new gsj_hostset();
    h->s = gsj::dedup(gro, rvo, xy, n, mo, mi);
    std::memcpy(moff, mo.data(), mo.size() * 8);
    std::memcpy(midx, mi.data(), mi.size() * 4);
    *n_distinct = h->s.n();
    *distinct = h;
  });
}

int gsj_subdivide_device(gsj_ctx* ctx, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv, gsj_hostset** out) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && gro && rvo && out, GSJ_E_ARG, "null argument");
    GSJ_REQUIRE(maxv >= 4, GSJ_E_ARG, "subdivide: maxv must be >= 4");
    gsj::check_offsets(gro, rvo, n);
    GSJ_REQUIRE(xy || rvo[gro[n]] == 0, GSJ_E_ARG, "null xy");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    auto* h = new gsj_hostset();
    try { h->s = gsj::subdivide_device(ctx->c, gro, rvo, xy, n, maxv); } catch (...) { delete h; throw; }
    *out = h;
  });
}

int gsj_dedup_device(gsj_ctx* ctx, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, gsj_hostset** distinct, uint64_t* moff,
                     uint32_t* midx, uint64_t* n_distinct) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && gro && rvo && distinct && moff && midx && n_distinct, GSJ_E_ARG, "null argument");
    gsj::check_offsets(gro, rvo, n);
    GSJ_REQUIRE(xy || rvo[gro[n]] == 0, GSJ_E_ARG, "null xy");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    std::vector<uint64_t> mo;
    std::vector<uint32_t> mi;
    auto* h = new gsj_hostset();
    try { h->s = gsj::dedup_device(ctx->c, gro, rvo, xy, n, mo, mi); } catch (...) { delete h; throw; }
    std::memcpy(moff, mo.data(), mo.size() * 8);
    std::memcpy(midx, mi.data(), mi.size() * 4);
    *n_distinct = h->s.n();
    *distinct = h;
  });
}

int gsj_copy_parse_device(gsj_ctx* ctx, const uint8_t* buf, uint64_t len, int header, int kind, gsj_hostset** polys, double** pts,
                          uint64_t* npts, char** ids, uint64_t** id_off, gsj_copy_stats* stats) {
  return guarded([&] {
    GSJ_REQUIRE(ctx && (buf || len == 0) && polys && pts && npts && ids && id_off && stats, GSJ_E_ARG, "null argument");
    *polys = nullptr;
    *pts = nullptr;
    *npts = 0;
    *ids = nullptr;
    *id_off = nullptr;
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    set_device(ctx->c);
    gsj::CopyLayer L = gsj::copy_parse_device(ctx->c, buf, len, header != 0, kind);
    const uint64_t nacc = L.points_layer ? L.npoints : L.npolys;
    GSJ_REQUIRE(L.id_off.size() == nacc + 1, GSJ_E_INTERNAL, "copy parse: id offsets do not match the accepted rows");
    auto take = [](size_t bytes) {
      void* p = std::malloc(std::max<size_t>(bytes, 1));
      if (!p) throw std::bad_alloc();
      return p;
    };
    std::unique_ptr<void, decltype(&std::free)> b_ids(take(L.ids.size()), &std::free), b_off(take((nacc + 1) * 8), &std::free),
        b_pts(nullptr, &std::free);
    std::memcpy(b_ids.get(), L.ids.data(), L.ids.size());
    std::memcpy(b_off.get(), L.id_off.data(), (nacc + 1) * 8);
    std::unique_ptr<gsj_hostset> set;
    if (L.points_layer) {
      b_pts.reset(take(L.pxy.size() * 8));
      std::memcpy(b_pts.get(), L.pxy.data(), L.pxy.size() * 8);
    } else {
      set.reset(new gsj_hostset());
      set->s = std::move(L.polys);
    }
    *stats = gsj_copy_stats{L.rows, L.nulls, L.bad_type, L.empty, L.npolys, L.npoints, L.consumed, L.done ? 1 : 0,
                            kind ? kind : (L.npolys ? 1 : (L.npoints ? 2 : 0)), L.scan_ms, L.emit_ms};
    *ids = static_cast<char*>(b_ids.release());
    *id_off = static_cast<uint64_t*>(b_off.release());
    if (L.points_layer) {
      *pts = static_cast<double*>(b_pts.release());
      *npts = L.npoints;
    } else {
      *polys = set.release();
    }
  });
}

int gsj_copy_write(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, const char* ids, const uint64_t* id_off, int srid,
                   uint8_t** out, uint64_t* len) {
  return guarded([&] {
    GSJ_REQUIRE(gro && rvo && out && len && (!ids || id_off), GSJ_E_ARG, "null argument");
    gsj::check_offsets(gro, rvo, n);
    GSJ_REQUIRE(xy || rvo[gro[n]] == 0, GSJ_E_ARG, "null xy");
    const std::string s = gsj::copy_write(gro, rvo, xy, n, ids, id_off, srid);
    void* p = std::malloc(s.size());
    if (!p) throw std::bad_alloc();
    std::memcpy(p, s.data(), s.size());
    *out = static_cast<uint8_t*>(p);
    *len = s.size();
  });
}

int gsj_generate(const char* kind, const char* spec, gsj_hostset** out) {
  return guarded([&] {
    GSJ_REQUIRE(kind && out, GSJ_E_ARG, "null argument");
    auto* h = new gsj_hostset();
    try { h->s = gsj::generate(kind, spec); } catch (...) { delete h; throw; }
    *out = h;
  });
}

int gsj_generate_points(const char* spec, double** xy, uint64_t* n) {
  return guarded([&] {
    GSJ_REQUIRE(xy && n, GSJ_E_ARG, "null argument");
    std::vector<double> v = gsj::generate_points(spec);
    *n = v.size() / 2;
    *xy = static_cast<double*>(std::malloc(std::max<size_t>(v.size(), 1) * 8));
    if (!*xy) throw std::bad_alloc();
    std::memcpy(*xy, v.data(), v.size() * 8);
  });
}

}  // extern "C"
