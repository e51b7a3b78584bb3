// gpujoind: the spatial join daemon on libgsj (reference: daemon/gpujoind.cu). Keeps named layers resident in HBM
// and serves the gpujoind wire protocol (protocol.h) over a unix socket and optionally TCP.
//
//   gpujoind [--socket PATH] [--tcp HOST:PORT] [--token-file FILE] [--lanes N] [--threads N] [--data-dir DIR]
//            [--device D] [--scratch-mb MB]
//
// Differences from the reference: requests do not serialise on a global GPU mutex; each connection's request goes to
// libgsj's scheduler (gsj_submit), whose lanes admit requests as they arrive and run them concurrently on the device
// (continuous batching). Rows are encoded and sent from the lane's callback, in parallel (OpenMP), as frames of up
// to ~1 MiB. A client that goes away cancels its request at the next sub-batch. LOAD reads layer files or
// generators (there is no Postgres here; see protocol.h).
#include "../../include/gsj.h"
#include "protocol.h"

#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <omp.h>
#include <poll.h>
#include <sys/socket.h>
#include <sys/stat.h>
#include <sys/un.h>
#include <unistd.h>

#include <algorithm>
#include <unordered_map>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <csignal>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;
double secs_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

std::string g_token;
std::string g_data_dir = ".";
int g_threads = 16;
gsj_ctx* g_ctx = nullptr;
gsj_sched* g_sched = nullptr;

void logf(const char* fmt, ...) {
  char ts[32];
  const time_t now = time(nullptr);
  struct tm tmv;
  localtime_r(&now, &tmv);
  strftime(ts, sizeof ts, "%Y-%m-%d %H:%M:%S", &tmv);
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "[%s] ", ts);
  vfprintf(stderr, fmt, ap);
  fputc('\n', stderr);
  va_end(ap);
}

void check(int rc, const char* what) {
  if (rc != GSJ_OK) throw std::runtime_error(std::string(what) + ": " + gsj_last_error());
}

// ------------------------------------------------------------------------------------------------- JSON
std::string jstr(const std::string& s) {
  std::string o = "\"";
  for (unsigned char c : s) {
    if (c == '"' || c == '\\') { o += '\\'; o += (char)c; }
    else if (c < 0x20) { char b[8]; snprintf(b, sizeof b, "\\u%04x", c); o += b; }
    else o += (char)c;
  }
  return o + "\"";
}
struct Json {
  std::string s = "{";
  bool first = true;
  Json& k(const char* key) { if (!first) s += ','; first = false; s += jstr(key) + ":"; return *this; }
  Json& v(const std::string& x) { s += jstr(x); return *this; }
  Json& n(double x) { char b[64]; snprintf(b, sizeof b, "%.6g", x); s += b; return *this; }
  Json& u(uint64_t x) { s += std::to_string(x); return *this; }
  Json& b(bool x) { s += x ? "true" : "false"; return *this; }
  Json& raw(const std::string& x) { s += x; return *this; }
  std::string done() const { return s + "}"; }
};

// ------------------------------------------------------------------------------------------------- layers
struct IdStore {   // row ids, as the extension sent them (text), packed
  std::string buf;
  std::vector<uint64_t> off{0};
  void add(const char* p, size_t n) { buf.append(p, n); off.push_back(buf.size()); }
  size_t size() const { return off.size() - 1; }
  const char* data(size_t i) const { return buf.data() + off[i]; }
  size_t len(size_t i) const { return off[i + 1] - off[i]; }
};

struct Layer {
  std::string name, kind, source;
  int subdiv = 0;
  IdStore ids;
  uint64_t rows = 0;
  bool deduped = false;
  std::vector<uint64_t> moff;   // distinct geometry -> members [moff[g], moff[g+1]) of midx
  std::vector<uint32_t> midx;
  uint64_t n_distinct = 0, gpu_geoms = 0, verts = 0, gpu_bytes = 0;
  double prep_s = 0, upload_s = 0;
  gsj_layer* gl = nullptr;
  template <class F> void members(uint32_t g, F&& f) const {
    if (!deduped) { f(g); return; }
    for (uint64_t k = moff[g]; k < moff[g + 1]; k++) f(midx[k]);
  }
  uint64_t msize(uint32_t g) const { return deduped ? moff[g + 1] - moff[g] : 1; }
  ~Layer() { if (gl) gsj_layer_free(gl); }   // refcounted: requests still running keep the device layer alive
  std::string json() const {
    return Json().k("layer").v(name).k("kind").v(kind).k("source").v(source).k("rows").u(rows).k("distinct").u(n_distinct)
        .k("gpu_geoms").u(gpu_geoms).k("vertices").u(verts).k("subdivide").u(subdiv).k("dedup").b(deduped)
        .k("gpu_mb").n(gpu_bytes / 1048576.0).k("prepare_s").n(prep_s).k("upload_s").n(upload_s).done();
  }
};

std::mutex g_reg;
std::map<std::string, std::shared_ptr<Layer>> g_layers;

std::shared_ptr<Layer> get_layer(const std::string& n) {
  std::lock_guard<std::mutex> g(g_reg);
  auto it = g_layers.find(n);
  return it == g_layers.end() ? nullptr : it->second;
}

std::string resolve(const std::string& p) { return (!p.empty() && p[0] == '/') ? p : g_data_dir + "/" + p; }

// ids from PATH.ids (one per line, as the reference's loader reads them), then row numbers for any rows left
void read_ids(const std::string& path, uint64_t n, IdStore& ids) {
  if (!path.empty()) {
    std::ifstream in(path + ".ids");
    std::string line;
    while (ids.size() < n && std::getline(in, line)) ids.add(line.data(), line.size());
  }
  for (uint64_t i = ids.size(); i < n; i++) { const std::string s = std::to_string(i); ids.add(s.data(), s.size()); }
}

struct HostsetGuard {
  gsj_hostset* h = nullptr;
  ~HostsetGuard() { if (h) gsj_hostset_free(h); }
};

// LOAD / LOAD_QUERY: source -> host set -> dedup -> subdivide -> device layer + BVH (reference build_layer,
// daemon/gpujoind.cu:100-149)
std::string build_layer(const std::string& name, const std::string& source, int subdiv, bool dedup) {
  auto t0 = Clock::now();
  auto L = std::make_shared<Layer>();
  L->name = name; L->source = source; L->subdiv = subdiv;
  HostsetGuard hs;
  std::vector<double> pts;
  bool is_point = false;
  if (source.rfind("gset:", 0) == 0) {
    const std::string path = resolve(source.substr(5));
    check(gsj_gset_read(path.c_str(), &hs.h), "reading layer file");
    uint64_t n, nr, nv;
    check(gsj_hostset_sizes(hs.h, &n, &nr, &nv), "layer sizes");
    read_ids(path, n, L->ids);
  } else if (source.rfind("pset:", 0) == 0) {
    const std::string path = resolve(source.substr(5));
    double* xy = nullptr;
    uint64_t n = 0;
    check(gsj_pset_read(path.c_str(), &xy, &n), "reading point file");
    pts.assign(xy, xy + 2 * n);
    gsj_free(xy);
    is_point = true;
    read_ids(path, n, L->ids);
  } else if (source.rfind("gen:", 0) == 0) {
    const size_t c = source.find(':', 4);
    if (c == std::string::npos) throw std::runtime_error("generator source must be gen:KIND:SPEC");
    check(gsj_generate(source.substr(4, c - 4).c_str(), source.substr(c + 1).c_str(), &hs.h), "generator");
    uint64_t n, nr, nv;
    check(gsj_hostset_sizes(hs.h, &n, &nr, &nv), "layer sizes");
    read_ids("", n, L->ids);
  } else if (source.rfind("genpts:", 0) == 0) {
    double* xy = nullptr;
    uint64_t n = 0;
    check(gsj_generate_points(source.substr(7).c_str(), &xy, &n), "point generator");
    pts.assign(xy, xy + 2 * n);
    gsj_free(xy);
    is_point = true;
    read_ids("", n, L->ids);
  } else {
    throw std::runtime_error("unknown layer source '" + source + "' (gset:PATH, pset:PATH, gen:KIND:SPEC, genpts:SPEC)");
  }
  L->rows = L->ids.size();
  if (is_point) {
    L->kind = "point";
    L->n_distinct = L->gpu_geoms = L->verts = L->rows;
    auto t1 = Clock::now();
    check(gsj_layer_create_points(g_ctx, pts.data(), L->rows, &L->gl), "upload");
    L->upload_s = secs_since(t1);
  } else {
    L->kind = "polygon";
    if (L->rows == 0) throw std::runtime_error("layer has no polygonal rows");
    const uint64_t *gro, *rvo;
    const double* xy;
    const uint32_t* parent;
    const uint8_t* syn;
    uint64_t n, nr, nv;
    check(gsj_hostset_arrays(hs.h, &gro, &rvo, &xy, &parent, &syn), "layer arrays");
    check(gsj_hostset_sizes(hs.h, &n, &nr, &nv), "layer sizes");
    HostsetGuard distinct, pieces;
    if (dedup) {
      L->moff.assign(n + 1, 0);
      L->midx.assign(n, 0);
      uint64_t nd = 0;
      check(gsj_dedup_device(g_ctx, gro, rvo, xy, n, &distinct.h, L->moff.data(), L->midx.data(), &nd), "dedup");   // on the GPU
      L->moff.resize(nd + 1);
      L->deduped = true;
      check(gsj_hostset_arrays(distinct.h, &gro, &rvo, &xy, &parent, &syn), "layer arrays");
      check(gsj_hostset_sizes(distinct.h, &n, &nr, &nv), "layer sizes");
    }
    L->n_distinct = n;
    if (subdiv > 0) {
      check(gsj_subdivide_device(g_ctx, gro, rvo, xy, n, subdiv, &pieces.h), "subdivide");   // on the GPU
      check(gsj_hostset_arrays(pieces.h, &gro, &rvo, &xy, &parent, &syn), "layer arrays");
      check(gsj_hostset_sizes(pieces.h, &n, &nr, &nv), "layer sizes");
    }
    L->gpu_geoms = n;
    L->verts = nv;
    L->prep_s = secs_since(t0);
    auto t1 = Clock::now();
    check(gsj_layer_create_polys(g_ctx, gro, rvo, xy, n, nr, nv, parent, syn, &L->gl), "upload");
    check(gsj_bvh_build(L->gl), "bvh");
    L->upload_s = secs_since(t1);
  }
  gsj_layer_info info;
  check(gsj_layer_info_get(L->gl, &info), "layer info");
  L->gpu_bytes = info.device_bytes;
  const std::string js = L->json();
  logf("load %s: ready %s", name.c_str(), js.c_str());
  std::shared_ptr<Layer> old;
  {
    std::lock_guard<std::mutex> g(g_reg);
    auto& slot = g_layers[name];
    old = slot;
    slot = L;
  }
  return js;
}

// ------------------------------------------------------------------------------------------------- io
struct Conn {
  int fd;
  volatile int broken = 0;      // set when a send fails; also the request's cancel flag (polled between sub-batches)
  bool streaming = false;       // frames started: errors go out as GSP_FRAME_ERR
  std::mutex wmu;
  bool read_full(void* p, size_t n) {
    char* c = (char*)p;
    while (n) {
      const ssize_t r = ::read(fd, c, n);
      if (r <= 0) { if (r < 0 && errno == EINTR) continue; return false; }
      c += r; n -= (size_t)r;
    }
    return true;
  }
  bool write_full(const void* p, size_t n) {
    const char* c = (const char*)p;
    while (n && !broken) {
      const ssize_t r = ::send(fd, c, n, MSG_NOSIGNAL);
      if (r <= 0) { if (r < 0 && errno == EINTR) continue; broken = 1; return false; }
      c += r; n -= (size_t)r;
    }
    return !broken;
  }
  bool u32(uint32_t v) { return write_full(&v, 4); }
  bool str(const std::string& s) { return u32((uint32_t)s.size()) && write_full(s.data(), s.size()); }
  void ok_json(const std::string& j) { u32(GSP_OK); str(j); }
  void error(const std::string& m) { u32(GSP_ERR); str(m); }
};

// frames of complete rows, closed at ~1 MiB (reference FrameBuf, daemon/gpujoind.cu:177-187)
struct FrameBuf {
  std::string out;
  size_t fstart = 0;
  uint32_t cnt = 0;
  uint64_t rows = 0;
  void begin() { fstart = out.size(); out.append(4, '\0'); cnt = 0; }
  void id(const char* p, size_t l) { const uint16_t L = (uint16_t)std::min<size_t>(l, 65535); out.append((const char*)&L, 2); out.append(p, L); }
  void f64(double x) { out.append((const char*)&x, 8); }
  void u32(uint32_t x) { out.append((const char*)&x, 4); }
  void row_done() { cnt++; rows++; if (out.size() - fstart >= (1u << 20)) { close(); begin(); } }
  void close() { memcpy(&out[fstart], &cnt, 4); }
  void finish() { if (cnt) close(); else out.resize(fstart); }
};

// encode items [0, n) in parallel slices (enc(i, frame)), send the slices in order; returns rows sent
template <class Enc>
uint64_t send_parallel(Conn& c, size_t n, Enc&& enc) {
  if (!n || c.broken) return 0;
  const int T = std::max(1, g_threads);
  const size_t nch = std::min<size_t>(n, (size_t)T * 4);
  std::vector<FrameBuf> fb(nch);
#pragma omp parallel for schedule(dynamic, 1) num_threads(T)
  for (int64_t k = 0; k < (int64_t)nch; k++) {
    const size_t i0 = n * k / nch, i1 = n * (k + 1) / nch;
    fb[k].out.reserve((i1 - i0) * 64);
    fb[k].begin();
    for (size_t i = i0; i < i1; i++) enc(i, fb[k]);
    fb[k].finish();
  }
  uint64_t rows = 0;
  std::lock_guard<std::mutex> g(c.wmu);
  for (auto& f : fb) { if (!c.write_full(f.out.data(), f.out.size())) break; rows += f.rows; }
  return rows;
}
void end_stream(Conn& c, const std::string& stats) { c.u32(0); c.str(stats); }

// run one request on the scheduler and wait for it
gsj_join_stats run_request(gsj_request& r, Conn& c, double* run_s = nullptr) {
  r.opts.cancel = &c.broken;
  uint64_t t = 0;
  check(gsj_submit(g_sched, &r, &t), "submit");
  gsj_join_stats st{};
  gsj_request_times tm{};
  const int rc = gsj_wait(g_sched, t, &st, &tm);
  if (rc == GSJ_E_CANCEL && c.broken) return st;
  check(rc, "request");
  if (run_s) *run_s = tm.run_ms / 1e3;
  return st;
}

const double NaN = std::numeric_limits<double>::quiet_NaN();

// self-join row expansion shared by JOIN, RELATE and DWITHIN (reference daemon/gpujoind.cu:231-283): members of
// both distinct geometries; a self-join lists each unordered pair once on the device, so rows are mirrored ("all") or
// ordered ("distinct")
struct Expand {
  const Layer* A;
  const Layer* B;
  bool same, distinct;
  // f(ra, rb, mirrored)
  template <class F> void rows(uint32_t pa, uint32_t pb, F&& f) const {
    const bool diag = same && pa == pb;
    A->members(pa, [&](uint32_t ra) {
      B->members(pb, [&](uint32_t rb) {
        if (!same) { f(ra, rb, false); return; }
        if (distinct) {
          if (ra == rb || (diag && ra > rb)) return;
          if (ra < rb) f(ra, rb, false);
          else f(rb, ra, !diag);
          return;
        }
        f(ra, rb, false);
        if (!diag) f(rb, ra, true);
      });
    });
  }
  uint64_t count(uint32_t pa, uint32_t pb) const {
    const uint64_t ma = A->msize(pa), mb = B->msize(pb);
    if (!same) return ma * mb;
    if (pa == pb) return distinct ? ma * (ma - 1) / 2 : ma * ma;
    return distinct ? ma * mb : 2 * ma * mb;
  }
};

// ------------------------------------------------------------------------------------------------- ops
struct JoinSink {
  Conn* c;
  Expand ex;
  bool stats_only, want_area, points, swapped;
  std::atomic<uint64_t> total{0};
  double area_sum = 0, host_s = 0;
};

int join_rows_cb(void* u, const uint32_t* pairs, const double* area, const uint32_t*, size_t n) {
  JoinSink& s = *(JoinSink*)u;
  auto th = Clock::now();
  if (s.stats_only) {
    uint64_t t = 0;
    double as = 0;
    for (size_t i = 0; i < n; i++) {
      const uint64_t k = s.ex.count(pairs[2 * i], pairs[2 * i + 1]);
      t += k;
      if (area) as += area[i] * (double)k;
    }
    s.total += t;
    s.area_sum += as;
  } else {
    s.total += send_parallel(*s.c, n, [&](size_t i, FrameBuf& f) {
      const double ar = area ? area[i] : NaN;
      s.ex.rows(pairs[2 * i], pairs[2 * i + 1], [&](uint32_t ra, uint32_t rb, bool) {
        f.id(s.ex.A->ids.data(ra), s.ex.A->ids.len(ra)); f.id(s.ex.B->ids.data(rb), s.ex.B->ids.len(rb)); f.f64(ar); f.row_done();
      });
    });
  }
  s.host_s += secs_since(th);
  return s.c->broken ? 1 : 0;
}

int join_points_cb(void* u, const uint32_t* pairs, const uint8_t*, size_t n) {
  JoinSink& s = *(JoinSink*)u;
  auto th = Clock::now();
  // points are the probe side; rows are (a_id, b_id) in the request's order
  const Layer* P = s.swapped ? s.ex.B : s.ex.A;
  const Layer* G = s.swapped ? s.ex.A : s.ex.B;
  if (s.stats_only) {
    uint64_t t = 0;
    for (size_t i = 0; i < n; i++) t += G->msize(pairs[2 * i + 1]);
    s.total += t;
  } else {
    s.total += send_parallel(*s.c, n, [&](size_t i, FrameBuf& f) {
      G->members(pairs[2 * i + 1], [&](uint32_t rg) {
        const uint32_t rp = pairs[2 * i];
        if (s.swapped) { f.id(G->ids.data(rg), G->ids.len(rg)); f.id(P->ids.data(rp), P->ids.len(rp)); }
        else { f.id(P->ids.data(rp), P->ids.len(rp)); f.id(G->ids.data(rg), G->ids.len(rg)); }
        f.f64(s.want_area ? 0.0 : NaN);
        f.row_done();
      });
    });
  }
  s.host_s += secs_since(th);
  return s.c->broken ? 1 : 0;
}

struct LayerPair {
  std::shared_ptr<Layer> A, B;
  bool same = false, swapped = false;
};
LayerPair layer_pair(const std::string& a, const std::string& b, const char* what) {
  LayerPair lp;
  lp.A = get_layer(a);
  lp.B = get_layer(b);
  if (!lp.A) throw std::runtime_error("no such layer: " + a);
  if (!lp.B) throw std::runtime_error("no such layer: " + b);
  if (lp.A->kind == "point" && lp.B->kind == "point") throw std::runtime_error(std::string("point x point ") + what + " is not supported");
  lp.same = lp.A.get() == lp.B.get();
  if (lp.A->kind == "polygon" && lp.B->kind == "point") lp.swapped = true;   // run as point x polygon
  return lp;
}
bool self_pairs_distinct(const std::string& s) {
  if (s != "all" && s != "distinct") throw std::runtime_error("self_pairs must be 'all' or 'distinct'");
  return s == "distinct";
}

double to_double(const std::string& s) {
  char* e;
  const double v = strtod(s.c_str(), &e);
  if (s.empty() || *e || !std::isfinite(v)) throw std::runtime_error("not a number: " + s);
  return v;
}

// Optional trailing arguments after an op's own (additive: the unchanged extension never sends them): "stats"
// (stats-only counting), "distance" (DWITHIN: the geography distance in the f64 slot), "window=x0,y0,x1,y1" (probe
// only the first layer's geometries, or the points of a point join, whose box meets the rectangle; a self-join keeps
// the mirrored rows of the pairs found).
struct Trailing {
  bool stats = false, distance = false, has_window = false;
  double w[4] = {0, 0, 0, 0};
  void apply(gsj_join_opts& o) const {
    o.has_window = has_window;
    for (int k = 0; k < 4; k++) o.window[k] = w[k];
  }
};
Trailing trailing(const std::vector<std::string>& a, size_t from, bool allow_distance) {
  Trailing t;
  for (size_t i = from; i < a.size(); i++) {
    if (a[i] == "stats") t.stats = true;
    else if (a[i] == "distance" && allow_distance) t.distance = true;
    else if (a[i].compare(0, 7, "window=") == 0) {
      std::vector<std::string> f;
      size_t p = 7;
      for (size_t q; (q = a[i].find(',', p)) != std::string::npos; p = q + 1) f.push_back(a[i].substr(p, q - p));
      f.push_back(a[i].substr(p));
      if (f.size() != 4) throw std::runtime_error("window needs x0,y0,x1,y1");
      for (int k = 0; k < 4; k++) t.w[k] = to_double(f[k]);
      if (!(t.w[0] <= t.w[2] && t.w[1] <= t.w[3])) throw std::runtime_error("window needs x0 <= x1 and y0 <= y1");
      t.has_window = true;
    } else {
      throw std::runtime_error("unknown option: " + a[i]);
    }
  }
  return t;
}

// JOIN (reference op_join, daemon/gpujoind.cu:210-307)
void op_join(Conn& c, const std::vector<std::string>& a) {
  if (a.size() < 4) throw std::runtime_error("JOIN needs layer_a, layer_b, want_area, self_pairs[, stats][, window=x0,y0,x1,y1]");
  const Trailing tr = trailing(a, 4, false);
  const bool stats_only = tr.stats;
  LayerPair lp = layer_pair(a[0], a[1], "joins");
  const bool want_area = a[2] == "1" || a[2] == "t" || a[2] == "true";
  const bool distinct = self_pairs_distinct(a[3]);
  c.u32(GSP_OK);
  c.streaming = true;
  auto t0 = Clock::now();
  JoinSink s{&c, Expand{lp.A.get(), lp.B.get(), lp.same, distinct}, stats_only, want_area, false, lp.swapped};
  gsj_request r{};
  std::string stats;
  if (lp.A->kind == "point" || lp.swapped) {
    const Layer* P = lp.swapped ? lp.B.get() : lp.A.get();
    const Layer* G = lp.swapped ? lp.A.get() : lp.B.get();
    s.points = true;
    r.kind = GSJ_REQ_POINTS_IN_POLYGONS; r.P = P->gl; r.R = G->gl; r.points_cb = join_points_cb; r.user = &s;
    tr.apply(r.opts);
    const gsj_join_stats st = run_request(r, c);
    stats = Json().k("kind").v("point_in_polygon").k("candidates").u(st.candidates).k("gpu_hits").u(st.hits).k("rows").u(s.total)
        .k("traverse_ms").n(st.traverse_ms).k("refine_ms").n(st.refine_ms).k("host_encode_send_s").n(s.host_s).k("total_s").n(secs_since(t0))
        .k("stats_only").b(stats_only).done();
  } else {
    r.kind = GSJ_REQ_JOIN; r.P = lp.A->gl; r.R = lp.B->gl; r.opts.want_area = want_area; r.opts.upper = lp.same;
    r.rows_cb = join_rows_cb; r.user = &s;
    tr.apply(r.opts);
    const gsj_join_stats st = run_request(r, c);
    Json j;
    j.k("kind").v("polygon_polygon").k("candidates").u(st.candidates).k("gpu_hits").u(st.hits).k("rows").u(s.total).k("heavy").u(st.heavy)
        .k("traverse_ms").n(st.traverse_ms).k("refine_ms").n(st.refine_ms).k("compact_ms").n(st.compact_ms).k("d2h_ms").n(st.copy_ms)
        .k("host_encode_send_s").n(s.host_s).k("total_s").n(secs_since(t0)).k("stats_only").b(stats_only);
    if (stats_only && want_area) j.k("area_sum").n(s.area_sum);
    stats = j.done();
  }
  if (!c.broken) end_stream(c, stats);
  logf("join %s x %s area=%d: %s%s", a[0].c_str(), a[1].c_str(), (int)want_area, stats.c_str(), c.broken ? " (client went away)" : "");
}

int cluster_rows_cb(void* u, const uint32_t* pairs, const double*, const uint32_t*, size_t n) {
  auto* v = static_cast<std::vector<std::pair<uint32_t, uint32_t>>*>(u);
  for (size_t i = 0; i < n; i++) v->push_back({pairs[2 * i], pairs[2 * i + 1]});
  return 0;
}

// CLUSTER (reference op_cluster, daemon/gpujoind.cu:290-322): dense cluster numbers in order of first row. With the
// trailing window=x0,y0,x1,y1 (an extension): the clusters of the layer's geometries whose box meets the window, rows
// for their source rows only, numbered the same way.
void op_cluster(Conn& c, const std::vector<std::string>& a) {
  if (a.empty()) throw std::runtime_error("CLUSTER needs layer[, window=x0,y0,x1,y1]");
  const Trailing tr = trailing(a, 1, false);
  if (tr.stats) throw std::runtime_error("CLUSTER takes no stats option");
  auto L = get_layer(a[0]);
  if (!L) throw std::runtime_error("no such layer: " + a[0]);
  if (L->kind != "polygon") throw std::runtime_error("cluster needs a polygon layer");
  auto t0 = Clock::now();
  gsj_request r{};
  std::vector<uint64_t> srows;     // source rows sent, ascending
  std::vector<uint32_t> row_root;  // their roots
  gsj_join_stats st{};
  if (!tr.has_window) {
    std::vector<uint32_t> root(std::max<uint64_t>(L->n_distinct, 1));
    r.kind = GSJ_REQ_CLUSTER; r.P = L->gl; r.root = root.data(); r.nroot = L->n_distinct;
    st = run_request(r, c);
    row_root.resize(L->rows);
    for (uint32_t gi = 0; gi < L->n_distinct; gi++) L->members(gi, [&](uint32_t rr) { row_root[rr] = root[gi]; });
    srows.resize(L->rows);
    for (uint64_t rr = 0; rr < L->rows; rr++) srows[rr] = rr;
  } else {
    std::vector<std::pair<uint32_t, uint32_t>> gr;   // (distinct geometry, root) of the window's geometries
    r.kind = GSJ_REQ_CLUSTER_ROWS; r.P = L->gl; r.rows_cb = cluster_rows_cb; r.user = &gr;
    tr.apply(r.opts);
    st = run_request(r, c);
    std::vector<std::pair<uint64_t, uint32_t>> rows;
    for (const auto& [gi, rt] : gr) L->members(gi, [&](uint32_t rr) { rows.push_back({rr, rt}); });
    std::sort(rows.begin(), rows.end());
    for (const auto& [rr, rt] : rows) { srows.push_back(rr); row_root.push_back(rt); }
  }
  std::vector<uint32_t> label(srows.size());
  uint32_t k = 0;
  if (!tr.has_window) {   // every geometry: a table over the roots
    std::vector<uint32_t> dense(std::max<uint64_t>(L->n_distinct, 1), UINT32_MAX);
    for (size_t i = 0; i < srows.size(); i++) {
      if (dense[row_root[i]] == UINT32_MAX) dense[row_root[i]] = k++;
      label[i] = dense[row_root[i]];
    }
  } else {                // a window's few geometries
    std::unordered_map<uint32_t, uint32_t> dense;
    for (size_t i = 0; i < srows.size(); i++) label[i] = dense.emplace(row_root[i], (uint32_t)dense.size()).first->second;
    k = (uint32_t)dense.size();
  }
  c.u32(GSP_OK);
  c.streaming = true;
  const uint64_t total = send_parallel(c, srows.size(), [&](size_t i, FrameBuf& f) {
    f.id(L->ids.data(srows[i]), L->ids.len(srows[i]));
    f.u32(label[i]);
    f.row_done();
  });
  const std::string stats = Json().k("kind").v("cluster_intersecting").k("rows").u(total).k("clusters").u(k).k("candidates").u(st.candidates)
      .k("edges").u(st.hits).k("traverse_ms").n(st.traverse_ms).k("refine_ms").n(st.refine_ms).k("total_s").n(secs_since(t0)).done();
  if (!c.broken) end_stream(c, stats);
  logf("cluster %s: %s", a[0].c_str(), stats.c_str());
}

// RELATE (reference op_relate, daemon/gpujoind.cu:325-389)
struct RelateSink {
  Conn* c;
  Expand ex;
  std::string pat;
  bool swapped;
  std::atomic<uint64_t> total{0};
  double host_s = 0;
  void put(FrameBuf& f, const Layer& LA, uint32_t ra, const Layer& LB, uint32_t rb, uint32_t code) const {
    if (!pat.empty() && !gsj_de9im_match(code, pat.c_str())) return;
    char m[9];
    gsj_de9im_str(code, m);
    f.id(LA.ids.data(ra), LA.ids.len(ra)); f.id(LB.ids.data(rb), LB.ids.len(rb)); f.out.append(m, 9); f.row_done();
  }
};
int relate_rows_cb(void* u, const uint32_t* pairs, const double*, const uint32_t* code, size_t n) {
  RelateSink& s = *(RelateSink*)u;
  auto th = Clock::now();
  s.total += send_parallel(*s.c, n, [&](size_t i, FrameBuf& f) {
    const uint32_t cd = code[i], ct = gsj_de9im_transpose(cd);
    s.ex.rows(pairs[2 * i], pairs[2 * i + 1], [&](uint32_t ra, uint32_t rb, bool mirrored) { s.put(f, *s.ex.A, ra, *s.ex.B, rb, mirrored ? ct : cd); });
  });
  s.host_s += secs_since(th);
  return s.c->broken ? 1 : 0;
}
int relate_points_cb(void* u, const uint32_t* pairs, const uint8_t* where, size_t n) {
  RelateSink& s = *(RelateSink*)u;
  auto th = Clock::now();
  const Layer* P = s.swapped ? s.ex.B : s.ex.A;
  const Layer* G = s.swapped ? s.ex.A : s.ex.B;
  s.total += send_parallel(*s.c, n, [&](size_t i, FrameBuf& f) {
    const uint32_t code = where[i] == 2 ? GSJ_DE9IM_PT_BOUNDARY : GSJ_DE9IM_PT_INTERIOR;   // relate(point, polygon)
    G->members(pairs[2 * i + 1], [&](uint32_t rg) {
      if (s.swapped) s.put(f, *G, rg, *P, pairs[2 * i], gsj_de9im_transpose(code));
      else s.put(f, *P, pairs[2 * i], *G, rg, code);
    });
  });
  s.host_s += secs_since(th);
  return s.c->broken ? 1 : 0;
}
void op_relate(Conn& c, const std::vector<std::string>& a) {
  if (a.size() < 4) throw std::runtime_error("RELATE needs layer_a, layer_b, pattern, self_pairs[, window=x0,y0,x1,y1]");
  const Trailing tr = trailing(a, 4, false);
  if (tr.stats) throw std::runtime_error("RELATE has no stats mode");
  LayerPair lp = layer_pair(a[0], a[1], "relate");
  const std::string pat = a[2];
  if (!pat.empty()) {
    if (pat.size() != 9) throw std::runtime_error("DE-9IM pattern must have 9 characters");
    for (char ch : pat)
      if (!strchr("TtFf*012", ch)) throw std::runtime_error("bad DE-9IM pattern character '" + std::string(1, ch) + "'");
  }
  const bool distinct = self_pairs_distinct(a[3]);
  c.u32(GSP_OK);
  c.streaming = true;
  auto t0 = Clock::now();
  RelateSink s{&c, Expand{lp.A.get(), lp.B.get(), lp.same, distinct}, pat, lp.swapped};
  gsj_request r{};
  gsj_join_stats st{};
  if (lp.A->kind == "point" || lp.swapped) {
    r.kind = GSJ_REQ_POINTS_IN_POLYGONS; r.P = (lp.swapped ? lp.B : lp.A)->gl; r.R = (lp.swapped ? lp.A : lp.B)->gl;
    r.points_cb = relate_points_cb; r.user = &s;
    tr.apply(r.opts);
    st = run_request(r, c);
  } else {
    r.kind = GSJ_REQ_JOIN; r.P = lp.A->gl; r.R = lp.B->gl; r.opts.want_relate = 1; r.opts.upper = lp.same;
    r.rows_cb = relate_rows_cb; r.user = &s;
    tr.apply(r.opts);
    st = run_request(r, c);
  }
  const std::string stats = Json().k("kind").v("relate").k("pattern").v(pat).k("candidates").u(st.candidates).k("gpu_hits").u(st.hits)
      .k("rows").u(s.total).k("host_encode_send_s").n(s.host_s).k("total_s").n(secs_since(t0)).done();
  if (!c.broken) end_stream(c, stats);
  logf("relate %s x %s: %s%s", a[0].c_str(), a[1].c_str(), stats.c_str(), c.broken ? " (client went away)" : "");
}

// DWITHIN (reference op_dwithin, daemon/gpujoind.cu:396-477)
struct DistSink {
  Conn* c;
  Expand ex;
  bool stats_only, points, swapped, with_distance;
  std::atomic<uint64_t> total{0};
  double host_s = 0;
};
int dist_rows_cb(void* u, const uint32_t* pairs, const double* dist, size_t n) {
  DistSink& s = *(DistSink*)u;
  auto th = Clock::now();
  if (s.points) {
    const Layer* P = s.swapped ? s.ex.B : s.ex.A;
    const Layer* G = s.swapped ? s.ex.A : s.ex.B;
    if (s.stats_only) {
      uint64_t t = 0;
      for (size_t i = 0; i < n; i++) t += G->msize(pairs[2 * i + 1]);
      s.total += t;
    } else {
      s.total += send_parallel(*s.c, n, [&](size_t i, FrameBuf& f) {
        const double v = (s.with_distance && dist) ? dist[i] : NaN;
        G->members(pairs[2 * i + 1], [&](uint32_t rg) {
          const uint32_t rp = pairs[2 * i];
          if (s.swapped) { f.id(G->ids.data(rg), G->ids.len(rg)); f.id(P->ids.data(rp), P->ids.len(rp)); }
          else { f.id(P->ids.data(rp), P->ids.len(rp)); f.id(G->ids.data(rg), G->ids.len(rg)); }
          f.f64(v);
          f.row_done();
        });
      });
    }
  } else if (s.stats_only) {
    uint64_t t = 0;
    for (size_t i = 0; i < n; i++) t += s.ex.count(pairs[2 * i], pairs[2 * i + 1]);
    s.total += t;
  } else {
    s.total += send_parallel(*s.c, n, [&](size_t i, FrameBuf& f) {
      const double v = (s.with_distance && dist) ? dist[i] : NaN;
      s.ex.rows(pairs[2 * i], pairs[2 * i + 1], [&](uint32_t ra, uint32_t rb, bool) {
        f.id(s.ex.A->ids.data(ra), s.ex.A->ids.len(ra)); f.id(s.ex.B->ids.data(rb), s.ex.B->ids.len(rb)); f.f64(v); f.row_done();
      });
    });
  }
  s.host_s += secs_since(th);
  return s.c->broken ? 1 : 0;
}
void op_dwithin(Conn& c, const std::vector<std::string>& a) {
  if (a.size() < 6) throw std::runtime_error("DWITHIN needs layer_a, layer_b, mode, dist_a, dist_b, self_pairs[, stats|distance][, window=x0,y0,x1,y1]");
  const Trailing tr = trailing(a, 6, true);
  const bool stats_only = tr.stats, with_distance = tr.distance;
  LayerPair lp = layer_pair(a[0], a[1], "distance joins");
  const std::string mode = a[2];
  const double da = to_double(a[3]), db = to_double(a[4]);
  const bool distinct = self_pairs_distinct(a[5]);
  bool erode = false;
  double d = 0, ra = 0, rb = 0;
  if (mode == "dwithin") {
    if (da < 0) throw std::runtime_error("distance must be >= 0");
    d = da;
  } else if (mode == "buffer") {
    if (da <= 0 && db <= 0 && (da < 0 || db < 0)) { erode = true; ra = -da; rb = -db; }
    else if (da >= 0 && db >= 0) d = da + db;
    else throw std::runtime_error("buffers of opposite sign are not supported");
  } else {
    throw std::runtime_error("mode must be 'dwithin' or 'buffer'");
  }
  if (lp.swapped) std::swap(ra, rb);
  const bool points = lp.A->kind == "point" || lp.swapped;
  if (points && erode && ra > 0) throw std::runtime_error("a point layer cannot have a negative buffer");
  c.u32(GSP_OK);
  c.streaming = true;
  auto t0 = Clock::now();
  DistSink s{&c, Expand{lp.A.get(), lp.B.get(), lp.same, distinct}, stats_only, points, lp.swapped, with_distance && !erode};
  gsj_request r{};
  r.dist_cb = dist_rows_cb;
  r.user = &s;
  std::vector<uint8_t> need;
  if (points) {
    r.P = (lp.swapped ? lp.B : lp.A)->gl;
    r.R = (lp.swapped ? lp.A : lp.B)->gl;
    r.kind = erode ? GSJ_REQ_POINTS_INTERSECTS_BUFFERED : GSJ_REQ_POINTS_DWITHIN;
    r.d = d; r.rb = rb; r.report_distance = s.with_distance;
  } else {
    r.P = lp.A->gl; r.R = lp.B->gl;
    r.opts.upper = lp.same; r.same = lp.same;
    if (erode) {
      r.kind = GSJ_REQ_INTERSECTS_BUFFERED; r.ra = ra; r.rb = rb;
      if (lp.same && distinct) {   // a geometry's pair with itself only matters for groups of identical rows
        need.assign(std::max<uint64_t>(lp.A->n_distinct, 1), 0);
        if (lp.A->deduped)
          for (uint64_t gi = 0; gi < lp.A->n_distinct; gi++) need[gi] = lp.A->moff[gi + 1] - lp.A->moff[gi] > 1;
        r.need_self = need.data();
      }
    } else {
      r.kind = GSJ_REQ_DWITHIN; r.d = d; r.report_distance = s.with_distance;
    }
  }
  tr.apply(r.opts);
  const gsj_join_stats st = run_request(r, c);
  const std::string stats = Json().k("kind").v(erode ? "buffer_intersects" : "dwithin").k("distance_m").n(d)
      .k("erode_a_m").n(lp.swapped ? rb : ra).k("erode_b_m").n(lp.swapped ? ra : rb).k("candidates").u(st.candidates).k("gpu_hits").u(st.hits)
      .k("rows").u(s.total).k("heavy").u(st.heavy).k("traverse_ms").n(st.traverse_ms).k("refine_ms").n(st.refine_ms)
      .k("compact_ms").n(st.compact_ms).k("d2h_ms").n(st.copy_ms).k("host_encode_send_s").n(s.host_s).k("total_s").n(secs_since(t0))
      .k("stats_only").b(stats_only).k("distance").b(s.with_distance).done();
  if (!c.broken) end_stream(c, stats);
  logf("dwithin %s x %s %s %s %s: %s%s", a[0].c_str(), a[1].c_str(), mode.c_str(), a[3].c_str(), a[4].c_str(), stats.c_str(),
       c.broken ? " (client went away)" : "");
}

std::string op_status() {
  static std::mutex mu;   // the context's status string is rewritten by every call
  std::lock_guard<std::mutex> lk(mu);
  std::string ls = "[";
  {
    std::lock_guard<std::mutex> r(g_reg);
    bool f = true;
    for (auto& kv : g_layers) { if (!f) ls += ","; f = false; ls += jstr(kv.first); }
  }
  std::string dev = gsj_ctx_status_json(g_ctx);   // {"device":..,"arch":..,"cu":..,"mem_total_mb":..,"mem_free_mb":..}
  dev.pop_back();
  return dev + ",\"layers\":" + ls + "]}";
}

std::string op_list() {
  std::string s = "[";
  std::lock_guard<std::mutex> r(g_reg);
  bool f = true;
  for (auto& kv : g_layers) { if (!f) s += ","; f = false; s += kv.second->json(); }
  return s + "]";
}

int to_int(const std::string& s, int dflt) {
  if (s.empty()) return dflt;
  char* e;
  const long v = strtol(s.c_str(), &e, 10);
  if (*e) throw std::runtime_error("not an integer: " + s);
  return (int)v;
}
bool to_bool(const std::string& s, bool dflt) { return s.empty() ? dflt : (s == "1" || s == "t" || s == "true" || s == "on"); }

// one connection, one request (reference handle, daemon/gpujoind.cu:516-566)
void handle(int fd) {
  Conn c{fd};
  try {
    uint32_t hdr[3];
    if (!c.read_full(hdr, 12) || hdr[0] != GSP_MAGIC || hdr[2] > 64) { close(fd); return; }
    std::vector<std::string> args(hdr[2]);
    for (auto& s : args) {
      uint32_t l;
      if (!c.read_full(&l, 4) || l > (64u << 20)) { close(fd); return; }
      s.resize(l);
      if (l && !c.read_full(&s[0], l)) { close(fd); return; }
    }
    {   // args[0] is the token: constant-time compare, then drop it
      const std::string tok = args.empty() ? std::string() : args[0];
      unsigned diff = tok.size() != g_token.size();
      for (size_t i = 0; i < g_token.size(); i++) diff |= (unsigned)(uint8_t)g_token[i] ^ (unsigned)(uint8_t)(i < tok.size() ? tok[i] : 0);
      if (diff) { logf("rejected request: bad token"); c.error("authentication failed"); close(fd); return; }
      if (!args.empty()) args.erase(args.begin());
    }
    switch (hdr[1]) {
      case GSP_OP_LOAD:
        if (args.size() < 7) throw std::runtime_error("LOAD needs 7 args");
        c.ok_json(build_layer(args[0], args[1], to_int(args[5], 0), to_bool(args[6], true)));
        break;
      case GSP_OP_LOAD_QUERY:
        if (args.size() < 4) throw std::runtime_error("LOAD_QUERY needs 4 args");
        c.ok_json(build_layer(args[0], args[1], to_int(args[2], 0), to_bool(args[3], true)));
        break;
      case GSP_OP_DROP: {
        if (args.empty()) throw std::runtime_error("DROP needs layer");
        std::shared_ptr<Layer> old;
        {
          std::lock_guard<std::mutex> g(g_reg);
          auto it = g_layers.find(args[0]);
          if (it != g_layers.end()) { old = it->second; g_layers.erase(it); }
        }
        c.ok_json(old ? "true" : "false");
        break;
      }
      case GSP_OP_LIST: c.ok_json(op_list()); break;
      case GSP_OP_STATUS: c.ok_json(op_status()); break;
      case GSP_OP_JOIN: op_join(c, args); break;
      case GSP_OP_CLUSTER: op_cluster(c, args); break;
      case GSP_OP_RELATE: op_relate(c, args); break;
      case GSP_OP_DWITHIN: op_dwithin(c, args); break;
      default: c.error("unknown op");
    }
  } catch (const std::exception& e) {
    logf("error: %s", e.what());
    if (c.streaming) { c.u32(GSP_FRAME_ERR); c.str(e.what()); }
    else c.error(e.what());
  }
  close(fd);
}

int listen_unix(const std::string& path) {
  const int fd = socket(AF_UNIX, SOCK_STREAM, 0);
  sockaddr_un a{};
  a.sun_family = AF_UNIX;
  strncpy(a.sun_path, path.c_str(), sizeof(a.sun_path) - 1);
  unlink(path.c_str());
  if (fd < 0 || bind(fd, (sockaddr*)&a, sizeof a) || listen(fd, 64)) throw std::runtime_error("unix socket " + path + ": " + strerror(errno));
  chmod(path.c_str(), 0666);
  return fd;
}
int listen_tcp(const std::string& hp) {
  const auto pos = hp.rfind(':');
  if (pos == std::string::npos) throw std::runtime_error("tcp address must be HOST:PORT");
  const std::string host = hp.substr(0, pos);
  const int port = atoi(hp.c_str() + pos + 1);
  const int fd = socket(AF_INET, SOCK_STREAM, 0), one = 1;
  setsockopt(fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
  sockaddr_in a{};
  a.sin_family = AF_INET;
  a.sin_port = htons((uint16_t)port);
  if (inet_pton(AF_INET, host.c_str(), &a.sin_addr) != 1) throw std::runtime_error("bad tcp address " + hp);
  if (fd < 0 || bind(fd, (sockaddr*)&a, sizeof a) || listen(fd, 64)) throw std::runtime_error("tcp " + hp + ": " + strerror(errno));
  return fd;
}

}  // namespace

int main(int argc, char** argv) {
  std::string sock = "/tmp/gpujoind.sock", tcp;
  int lanes = 4, device = 0;
  uint64_t scratch_mb = 0;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--socket" && i + 1 < argc) sock = argv[++i];
    else if (a == "--tcp" && i + 1 < argc) tcp = argv[++i];
    else if (a == "--threads" && i + 1 < argc) g_threads = atoi(argv[++i]);
    else if (a == "--lanes" && i + 1 < argc) lanes = atoi(argv[++i]);
    else if (a == "--device" && i + 1 < argc) device = atoi(argv[++i]);
    else if (a == "--scratch-mb" && i + 1 < argc) scratch_mb = strtoull(argv[++i], nullptr, 10);
    else if (a == "--data-dir" && i + 1 < argc) g_data_dir = argv[++i];
    else if (a == "--token-file" && i + 1 < argc) {
      std::ifstream tf(argv[++i]);
      std::getline(tf, g_token);
      if (g_token.empty()) { fprintf(stderr, "empty token file\n"); return 1; }
    } else {
      fprintf(stderr, "usage: gpujoind [--socket PATH] [--tcp HOST:PORT] [--token-file FILE] [--lanes N] [--threads N] [--data-dir DIR] "
                      "[--device D] [--scratch-mb MB]\n");
      return 1;
    }
  }
  if (!tcp.empty() && g_token.empty()) { fprintf(stderr, "--tcp requires --token-file\n"); return 1; }
  signal(SIGPIPE, SIG_IGN);
  try {
    check(gsj_ctx_create(device, &g_ctx), "device");
    check(gsj_sched_create(g_ctx, lanes, scratch_mb << 20, &g_sched), "scheduler");
    logf("gpujoind on %s, %d lanes", gsj_ctx_status_json(g_ctx), lanes);
    std::vector<pollfd> fds;
    fds.push_back({listen_unix(sock), POLLIN, 0});
    logf("listening on unix:%s", sock.c_str());
    if (!tcp.empty()) { fds.push_back({listen_tcp(tcp), POLLIN, 0}); logf("listening on tcp:%s", tcp.c_str()); }
    for (;;) {
      if (poll(fds.data(), fds.size(), -1) < 0) {
        if (errno == EINTR) continue;
        perror("poll");
        return 1;
      }
      for (auto& f : fds)
        if (f.revents & POLLIN) {
          const int cfd = accept(f.fd, nullptr, nullptr);
          if (cfd < 0) continue;
          const int one = 1;
          setsockopt(cfd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof one);
          std::thread(handle, cfd).detach();
        }
    }
  } catch (const std::exception& e) {
    fprintf(stderr, "gpujoind: %s\n", e.what());
    return 1;
  }
}
