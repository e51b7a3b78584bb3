// gpujoind: the spatial join daemon on libgsj (reference: daemon/gpujoind.cu). Keeps named layers resident in HBM
// and serves the gpujoind wire protocol (protocol.h) over a unix socket and optionally TCP.
//
//   gpujoind [--socket PATH] [--socket-mode OCTAL] [--tcp HOST:PORT] [--token-file FILE] [--lanes N] [--threads N]
//            [--data-dir DIR] [--device D] [--scratch-mb MB] [--max-rows N] [--batch N]
//
// Differences from the reference: requests do not serialise on a global GPU mutex; each connection's request goes to
// libgsj's scheduler (gsj_submit), whose lanes admit requests as they arrive and run them concurrently on the device
// (continuous batching). Rows are encoded from the lane's callback on several threads (OpenMP) and leave as frames of
// at most 1 MiB of rows, written straight from the encoded chunks. A client that goes away cancels its request at the
// next sub-batch; --max-rows ends a stream that would exceed it with a GSP_FRAME_ERR frame. LOAD reads layer files
// under --data-dir (paths that resolve outside it are refused) or generators (there is no Postgres here; see
// protocol.h).
#include "../../include/gsj.h"
#include "protocol.h"

#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <omp.h>
#include <poll.h>
#include <sys/socket.h>
#include <sys/stat.h>
#include <sys/uio.h>
#include <sys/un.h>
#include <unistd.h>

#include <algorithm>
#include <unordered_map>
#include <atomic>
#include <cerrno>
#include <climits>
#include <condition_variable>
#include <deque>
#include <functional>
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
std::string g_data_real;   // its real path (set at start)
int g_threads = 16;
uint64_t g_max_rows = 0;   // 0: no limit on the rows of one stream
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
  std::vector<uint32_t> self_needed;   // distinct geometries of more than one row, ascending (built once, at LOAD)
  uint64_t n_distinct = 0, gpu_geoms = 0, verts = 0, gpu_bytes = 0;
  uint64_t copy_rows = 0, skipped_null = 0, skipped_type = 0, skipped_empty = 0;   // copy: sources (the loader's counters)
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
        .k("gpu_mb").n(gpu_bytes / 1048576.0).k("prepare_s").n(prep_s).k("upload_s").n(upload_s)
        .k("stream_rows").u(copy_rows).k("skipped_null").u(skipped_null).k("skipped_type").u(skipped_type).k("skipped_empty").u(skipped_empty)
        .done();
  }
};

std::mutex g_reg;
std::map<std::string, std::shared_ptr<Layer>> g_layers;

std::shared_ptr<Layer> get_layer(const std::string& n) {
  std::lock_guard<std::mutex> g(g_reg);
  auto it = g_layers.find(n);
  return it == g_layers.end() ? nullptr : it->second;
}

// a layer file named by a LOAD: relative to --data-dir, and after resolving links and "..", still inside it
std::string resolve(const std::string& p) {
  const std::string full = (!p.empty() && p[0] == '/') ? p : g_data_dir + "/" + p;
  char real[PATH_MAX];
  if (!::realpath(full.c_str(), real)) throw std::runtime_error("cannot open " + p + ": " + strerror(errno));
  const std::string r = real;
  if (r.size() <= g_data_real.size() || r.compare(0, g_data_real.size(), g_data_real) != 0 || r[g_data_real.size()] != '/')
    throw std::runtime_error("layer file " + p + " is outside the data directory");
  return r;
}

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

// a layer read from a binary COPY file (the stream the reference's loader reads from Postgres, daemon/loader.h:176-279)
struct CopyRows {
  std::vector<uint64_t> gro{0}, rvo{0};
  std::vector<double> xy, pts;
  uint64_t rows = 0, nulls = 0, bad_type = 0, empty = 0;
  int kind = 0;   // 1 polygons, 2 points
};

// the file in pieces of 256 MiB, each parsed on the device; the rest of a piece (an incomplete row) and the layer
// kind carry over to the next, so the result is that of one parse of the whole stream
void read_copy_file(const std::string& path, CopyRows& out, IdStore& ids) {
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path.c_str(), "rb"), &std::fclose);
  if (!f) throw std::runtime_error("cannot open " + path + ": " + strerror(errno));
  constexpr size_t PIECE = 256u << 20;
  std::vector<uint8_t> buf;
  size_t have = 0;
  bool header = true, eof = false;
  while (!eof) {
    buf.resize(have + PIECE);
    const size_t got = std::fread(buf.data() + have, 1, PIECE, f.get());
    have += got;
    eof = got < PIECE;
    HostsetGuard ph;
    double* pp = nullptr;
    char* pid = nullptr;
    uint64_t* poff = nullptr;
    uint64_t np = 0;
    gsj_copy_stats st{};
    check(gsj_copy_parse_device(g_ctx, buf.data(), have, header ? 1 : 0, out.kind, &ph.h, &pp, &np, &pid, &poff, &st), "parsing the COPY stream");
    std::unique_ptr<double, void (*)(void*)> g_pts(pp, &gsj_free);
    std::unique_ptr<char, void (*)(void*)> g_ids(pid, &gsj_free);
    std::unique_ptr<uint64_t, void (*)(void*)> g_off(poff, &gsj_free);
    if (header && st.consumed == 0) {
      if (eof) throw std::runtime_error("COPY stream " + path + " ends inside its header");
      continue;   // the header is not complete yet
    }
    header = false;
    out.kind = st.kind;
    out.rows += st.rows;
    out.nulls += st.skipped_null;
    out.bad_type += st.skipped_type;
    out.empty += st.skipped_empty;
    const uint64_t nacc = st.polys + st.points;
    for (uint64_t k = 0; k < nacc; k++) ids.add(pid + poff[k], poff[k + 1] - poff[k]);
    if (pp) out.pts.insert(out.pts.end(), pp, pp + 2 * np);
    if (ph.h && st.polys) {
      const uint64_t *gro, *rvo;
      const double* xy;
      const uint32_t* parent;
      const uint8_t* syn;
      uint64_t n, nr, nv;
      check(gsj_hostset_arrays(ph.h, &gro, &rvo, &xy, &parent, &syn), "layer arrays");
      check(gsj_hostset_sizes(ph.h, &n, &nr, &nv), "layer sizes");
      const uint64_t r0 = out.gro.back(), v0 = out.rvo.back();
      for (uint64_t g = 1; g <= n; g++) out.gro.push_back(r0 + gro[g]);
      for (uint64_t r = 1; r <= nr; r++) out.rvo.push_back(v0 + rvo[r]);
      out.xy.insert(out.xy.end(), xy, xy + 2 * nv);
    }
    std::memmove(buf.data(), buf.data() + st.consumed, have - st.consumed);
    have -= st.consumed;
    if (st.done) break;
  }
}

// LOAD / LOAD_QUERY: source -> host set -> dedup -> subdivide -> device layer + BVH (reference build_layer,
// daemon/gpujoind.cu:100-149)
std::string build_layer(const std::string& name, const std::string& source, int subdiv, bool dedup) {
  auto t0 = Clock::now();
  auto L = std::make_shared<Layer>();
  L->name = name; L->source = source; L->subdiv = subdiv;
  HostsetGuard hs;
  std::vector<double> pts;
  bool is_point = false;
  CopyRows cr;
  const bool from_copy = source.rfind("copy:", 0) == 0;
  if (from_copy) {
    read_copy_file(resolve(source.substr(5)), cr, L->ids);
    L->skipped_null = cr.nulls;
    L->skipped_type = cr.bad_type;
    L->skipped_empty = cr.empty;
    L->copy_rows = cr.rows;
    if (cr.kind == 2) {
      pts.swap(cr.pts);
      is_point = true;
    }
  } else if (source.rfind("gset:", 0) == 0) {
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
    throw std::runtime_error("unknown layer source '" + source + "' (gset:PATH, pset:PATH, copy:PATH, gen:KIND:SPEC, genpts:SPEC)");
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
    if (from_copy) {
      gro = cr.gro.data();
      rvo = cr.rvo.data();
      xy = cr.xy.data();
      parent = nullptr;
      syn = nullptr;
      n = cr.gro.size() - 1;
      nr = cr.rvo.size() - 1;
      nv = cr.xy.size() / 2;
    } else {
      check(gsj_hostset_arrays(hs.h, &gro, &rvo, &xy, &parent, &syn), "layer arrays");
      check(gsj_hostset_sizes(hs.h, &n, &nr, &nv), "layer sizes");
    }
    HostsetGuard distinct, pieces;
    if (dedup) {
      L->moff.assign(n + 1, 0);
      L->midx.assign(n, 0);
      uint64_t nd = 0;
      check(gsj_dedup_device(g_ctx, gro, rvo, xy, n, &distinct.h, L->moff.data(), L->midx.data(), &nd), "dedup");   // on the GPU
      L->moff.resize(nd + 1);
      L->deduped = true;
      for (uint64_t gi = 0; gi < nd; gi++)
        if (L->moff[gi + 1] - L->moff[gi] > 1) L->self_needed.push_back((uint32_t)gi);
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


// ------------------------------------------------------------------------------------------------- connection
// One accepted connection. Reads wait at most IDLE_MS for each piece of the request, so a client that stalls
// mid-request releases its thread; `gone` is set when the peer stops reading or writing and doubles as the running
// request's cancel flag (the engine polls it between sub-batches).
constexpr int IDLE_MS = 120000;

class Peer {
 public:
  explicit Peer(int fd) : fd_(fd) {}
  ~Peer() { ::close(fd_); }
  Peer(const Peer&) = delete;
  Peer& operator=(const Peer&) = delete;
  volatile int gone = 0;

  bool read_exact(void* dst, size_t n) {
    auto* p = static_cast<char*>(dst);
    while (n > 0) {
      pollfd w{fd_, POLLIN, 0};
      const int pr = ::poll(&w, 1, IDLE_MS);
      if (pr < 0 && errno == EINTR) continue;
      if (pr <= 0) return false;   // idle too long, or poll failed
      const ssize_t got = ::recv(fd_, p, n, 0);
      if (got < 0 && errno == EINTR) continue;
      if (got <= 0) return false;
      p += got;
      n -= (size_t)got;
    }
    return true;
  }
  // every byte of the vectors, in order; on failure the peer is marked gone and later writes are dropped
  bool write_all(iovec* v, int cnt) {
    std::lock_guard<std::mutex> lk(wmu_);
    while (cnt > 0 && !gone) {
      msghdr m{};
      m.msg_iov = v;
      m.msg_iovlen = (size_t)cnt;
      const ssize_t put = ::sendmsg(fd_, &m, MSG_NOSIGNAL);
      if (put < 0 && errno == EINTR) continue;
      if (put <= 0) { gone = 1; break; }
      size_t left = (size_t)put;
      while (cnt > 0 && left >= v->iov_len) { left -= v->iov_len; v++; cnt--; }
      if (cnt > 0) { v->iov_base = static_cast<char*>(v->iov_base) + left; v->iov_len -= left; }
    }
    return !gone;
  }
  bool write_u32s(std::initializer_list<uint32_t> xs) {
    uint32_t buf[4];
    size_t k = 0;
    for (uint32_t x : xs) buf[k++] = x;
    iovec v{buf, k * 4};
    return write_all(&v, 1);
  }
  bool write_text(uint32_t lead, const std::string& s) {   // u32 lead, u32 length, bytes
    uint32_t hdr[2] = {lead, (uint32_t)s.size()};
    iovec v[2] = {{hdr, 8}, {const_cast<char*>(s.data()), s.size()}};
    return write_all(v, 2);
  }

 private:
  int fd_;
  std::mutex wmu_;
};

// ------------------------------------------------------------------------------------------------- request
constexpr uint32_t MAX_ARGS = 64;                 // protocol limits
constexpr uint32_t MAX_ARG_BYTES = 64u << 20;
constexpr uint64_t MAX_REQUEST_BYTES = 256ull << 20;

struct WireRequest {
  uint32_t op = 0;
  std::string token;
  std::vector<std::string> args;   // the op's arguments (the token removed)
};

// false for anything but a complete, well-formed request within the limits
bool read_request(Peer& p, WireRequest& q) {
  uint32_t head[3];
  if (!p.read_exact(head, sizeof head) || head[0] != GSP_MAGIC || head[2] == 0 || head[2] > MAX_ARGS) return false;
  q.op = head[1];
  std::vector<std::string> all(head[2]);
  uint64_t total = 0;
  for (std::string& a : all) {
    uint32_t len = 0;
    if (!p.read_exact(&len, 4) || len > MAX_ARG_BYTES || (total += len) > MAX_REQUEST_BYTES) return false;
    a.resize(len);
    if (len && !p.read_exact(&a[0], len)) return false;
  }
  q.token = std::move(all[0]);
  q.args.assign(std::make_move_iterator(all.begin() + 1), std::make_move_iterator(all.end()));
  return true;
}

// whether the client's token is the daemon's: the time taken does not depend on where the two first differ (all
// positions of the longer one are folded in)
bool token_matches(const std::string& got) {
  const size_t n = std::max(got.size(), g_token.size());
  uint8_t diff = got.size() == g_token.size() ? 0 : 1;
  for (size_t i = 0; i < n; i++) {
    const uint8_t a = i < got.size() ? (uint8_t)got[i] : 0, b = i < g_token.size() ? (uint8_t)g_token[i] : 0;
    diff = (uint8_t)(diff | (a ^ b));
  }
  return diff == 0;
}

// ------------------------------------------------------------------------------------------------- rows
// Rows of one callback batch, encoded on several threads into chunks: each chunk keeps its rows' bytes back to back
// and the end offset of every row, and the reply cuts frames from it at row boundaries (at most FRAME_BYTES of rows
// a frame) and hands each frame to the socket as two pieces, its count and a slice of the chunk.
constexpr size_t FRAME_BYTES = 1u << 20;

struct RowChunk {
  std::string bytes;
  std::vector<uint32_t> ends;
  void id(const char* p, size_t len) {
    const uint16_t l = (uint16_t)std::min<size_t>(len, 65535);   // ids longer than the u16 field are cut
    bytes.append(reinterpret_cast<const char*>(&l), 2);
    bytes.append(p, l);
  }
  void f64(double x) { bytes.append(reinterpret_cast<const char*>(&x), 8); }
  void u32(uint32_t x) { bytes.append(reinterpret_cast<const char*>(&x), 4); }
  void raw(const char* p, size_t len) { bytes.append(p, len); }
  void end_row() { ends.push_back((uint32_t)bytes.size()); }
};

// The response side of one request. JSON ops answer with json(); streaming ops open the stream, send row batches and
// close it with the stats; fail() answers GSP_ERR before the stream opened and a GSP_FRAME_ERR frame after. A row
// limit (--max-rows) stops a stream that would exceed it: the batch is not sent, the engine is asked to stop, and the
// op ends the stream with GSP_FRAME_ERR.
// While a request runs on the scheduler (between begin_async and pump), row batches encoded on the lane's thread are
// queued and the connection's own thread writes them (pump), so a lane never waits for a client to read: it goes on
// to its next request while the rows of this one drain.
class Reply {
 public:
  explicit Reply(Peer& p) : peer_(p) {}
  Peer& peer() { return peer_; }
  bool streaming() const { return streaming_; }
  bool over_limit() const { return over_limit_; }
  // nonzero: the engine should stop delivering rows (client gone, or row limit reached)
  int stop() const { return (peer_.gone || over_limit_) ? 1 : 0; }

  void json(const std::string& j) { peer_.write_text(GSP_OK, j); }
  void fail(const std::string& msg) { peer_.write_text(streaming_ ? GSP_FRAME_ERR : GSP_ERR, msg); }
  void open_stream() {
    peer_.write_u32s({GSP_OK});
    streaming_ = true;
  }
  void close_stream(const std::string& stats) {
    if (over_limit_) throw std::runtime_error("row limit of " + std::to_string(g_max_rows) + " rows exceeded");
    if (!peer_.gone) peer_.write_text(0, stats);
  }

  // encode items [0, n) with enc(i, chunk) on the encoding threads and send their rows in item order (queued for pump
  // while a request runs); returns the rows taken
  template <class Enc>
  uint64_t send_rows(size_t n, Enc&& enc) {
    if (n == 0 || stop()) return 0;
    const int T = std::max(1, g_threads);
    const size_t nchunk = std::min<size_t>(n, (size_t)T * 4);
    std::vector<RowChunk> ch(nchunk);
#pragma omp parallel for schedule(dynamic, 1) num_threads(T)
    for (int64_t k = 0; k < (int64_t)nchunk; k++) {
      const size_t lo = n * (size_t)k / nchunk, hi = n * (size_t)(k + 1) / nchunk;
      for (size_t i = lo; i < hi; i++) enc(i, ch[k]);
    }
    uint64_t rows = 0;
    for (const RowChunk& c : ch) rows += c.ends.size();
    if (g_max_rows && taken_ + rows > g_max_rows) { over_limit_ = true; return 0; }
    taken_ += rows;
    if (async_) {
      std::lock_guard<std::mutex> lk(qmu_);
      queue_.push_back(std::move(ch));
      qcv_.notify_one();
    } else {
      write_chunks(ch);
    }
    return rows;
  }

  // the request about to be submitted streams its rows through the queue until pump returns
  void begin_async() {
    async_ = true;
    done_ = false;
  }
  // gsj_request::done_cb: the request has finished, no row batch follows (notified under the lock: pump may return
  // and the reply go away as soon as it is released)
  static void request_done(void* self) {
    Reply& r = *static_cast<Reply*>(self);
    std::lock_guard<std::mutex> lk(r.qmu_);
    r.done_ = true;
    r.qcv_.notify_one();
  }
  // on the connection's thread: write queued batches as they come until the request has finished and all are out
  void pump() {
    for (;;) {
      std::vector<RowChunk> ch;
      {
        std::unique_lock<std::mutex> lk(qmu_);
        qcv_.wait(lk, [&] { return !queue_.empty() || done_; });
        if (queue_.empty()) break;   // done, nothing left
        ch = std::move(queue_.front());
        queue_.pop_front();
      }
      write_chunks(ch);   // after the client went away: dropped
    }
    async_ = false;
  }

 private:
  // frames of at most FRAME_BYTES of rows cut from each chunk at row boundaries, each written as its count and a slice
  void write_chunks(std::vector<RowChunk>& ch) {
    for (RowChunk& c : ch) {
      size_t first = 0;   // first row of the next frame
      while (first < c.ends.size() && !peer_.gone) {
        const uint32_t base = first ? c.ends[first - 1] : 0;
        size_t last = first + 1;   // one past the frame's last row: as many rows as fit FRAME_BYTES (at least one)
        while (last < c.ends.size() && c.ends[last] - base <= FRAME_BYTES) last++;
        uint32_t count = (uint32_t)(last - first);
        iovec v[2] = {{&count, 4}, {&c.bytes[base], c.ends[last - 1] - base}};
        if (!peer_.write_all(v, 2)) break;
        first = last;
      }
    }
  }

  Peer& peer_;
  bool streaming_ = false, over_limit_ = false;
  uint64_t taken_ = 0;   // rows taken for this stream (the row limit's count), on the thread that calls send_rows
  bool async_ = false;   // set before the submit and cleared by pump on the connection's thread: the lane reads it
                         // only while the request runs, between the two
  std::mutex qmu_;
  std::condition_variable qcv_;
  std::deque<std::vector<RowChunk>> queue_;
  bool done_ = false;
};

// run one request on the scheduler, wait for it, and return its stats; a stop asked for by the reply (client gone,
// row limit) is not an error here
gsj_join_stats run_request(gsj_request& r, Reply& rep, double* run_s = nullptr) {
  r.opts.cancel = &rep.peer().gone;
  r.done_cb = &Reply::request_done;
  r.done_user = &rep;
  uint64_t t = 0;
  rep.begin_async();
  const int src = gsj_submit(g_sched, &r, &t);
  if (src != GSJ_OK) Reply::request_done(&rep);   // nothing will run: pump returns at once
  rep.pump();
  check(src, "submit");
  gsj_join_stats st{};
  gsj_request_times tm{};
  const int rc = gsj_wait(g_sched, t, &st, &tm);
  if (rc == GSJ_E_CANCEL && rep.stop()) return st;
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
  Reply* c;
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
    s.total += s.c->send_rows(n, [&](size_t i, RowChunk& f) {
      const double ar = area ? area[i] : NaN;
      s.ex.rows(pairs[2 * i], pairs[2 * i + 1], [&](uint32_t ra, uint32_t rb, bool) {
        f.id(s.ex.A->ids.data(ra), s.ex.A->ids.len(ra)); f.id(s.ex.B->ids.data(rb), s.ex.B->ids.len(rb)); f.f64(ar); f.end_row();
      });
    });
  }
  s.host_s += secs_since(th);
  return s.c->stop();
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
    s.total += s.c->send_rows(n, [&](size_t i, RowChunk& f) {
      G->members(pairs[2 * i + 1], [&](uint32_t rg) {
        const uint32_t rp = pairs[2 * i];
        if (s.swapped) { f.id(G->ids.data(rg), G->ids.len(rg)); f.id(P->ids.data(rp), P->ids.len(rp)); }
        else { f.id(P->ids.data(rp), P->ids.len(rp)); f.id(G->ids.data(rg), G->ids.len(rg)); }
        f.f64(s.want_area ? 0.0 : NaN);
        f.end_row();
      });
    });
  }
  s.host_s += secs_since(th);
  return s.c->stop();
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
void op_join(Reply& c, const std::vector<std::string>& a) {
  if (a.size() < 4) throw std::runtime_error("JOIN needs layer_a, layer_b, want_area, self_pairs[, stats][, window=x0,y0,x1,y1]");
  const Trailing tr = trailing(a, 4, false);
  const bool stats_only = tr.stats;
  LayerPair lp = layer_pair(a[0], a[1], "joins");
  const bool want_area = a[2] == "1" || a[2] == "t" || a[2] == "true";
  const bool distinct = self_pairs_distinct(a[3]);
  c.open_stream();
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
  c.close_stream(stats);
  logf("join %s x %s area=%d: %s%s", a[0].c_str(), a[1].c_str(), (int)want_area, stats.c_str(), c.peer().gone ? " (client went away)" : "");
}

int cluster_rows_cb(void* u, const uint32_t* pairs, const double*, const uint32_t*, size_t n) {
  auto* v = static_cast<std::vector<std::pair<uint32_t, uint32_t>>*>(u);
  for (size_t i = 0; i < n; i++) v->push_back({pairs[2 * i], pairs[2 * i + 1]});
  return 0;
}

// CLUSTER (reference op_cluster, daemon/gpujoind.cu:290-322): dense cluster numbers in order of first row. With the
// trailing window=x0,y0,x1,y1 (an extension): the clusters of the layer's geometries whose box meets the window, rows
// for their source rows only, numbered the same way.
void op_cluster(Reply& c, const std::vector<std::string>& a) {
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
  c.open_stream();
  const uint64_t total = c.send_rows(srows.size(), [&](size_t i, RowChunk& f) {
    f.id(L->ids.data(srows[i]), L->ids.len(srows[i]));
    f.u32(label[i]);
    f.end_row();
  });
  const std::string stats = Json().k("kind").v("cluster_intersecting").k("rows").u(total).k("clusters").u(k).k("candidates").u(st.candidates)
      .k("edges").u(st.hits).k("traverse_ms").n(st.traverse_ms).k("refine_ms").n(st.refine_ms).k("total_s").n(secs_since(t0)).done();
  c.close_stream(stats);
  logf("cluster %s: %s", a[0].c_str(), stats.c_str());
}

// RELATE (reference op_relate, daemon/gpujoind.cu:325-389)
struct RelateSink {
  Reply* c;
  Expand ex;
  std::string pat;
  bool swapped;
  std::atomic<uint64_t> total{0};
  double host_s = 0;
  void put(RowChunk& f, const Layer& LA, uint32_t ra, const Layer& LB, uint32_t rb, uint32_t code) const {
    if (!pat.empty() && !gsj_de9im_match(code, pat.c_str())) return;
    char m[9];
    gsj_de9im_str(code, m);
    f.id(LA.ids.data(ra), LA.ids.len(ra)); f.id(LB.ids.data(rb), LB.ids.len(rb)); f.raw(m, 9); f.end_row();
  }
};
int relate_rows_cb(void* u, const uint32_t* pairs, const double*, const uint32_t* code, size_t n) {
  RelateSink& s = *(RelateSink*)u;
  auto th = Clock::now();
  s.total += s.c->send_rows(n, [&](size_t i, RowChunk& f) {
    const uint32_t cd = code[i], ct = gsj_de9im_transpose(cd);
    s.ex.rows(pairs[2 * i], pairs[2 * i + 1], [&](uint32_t ra, uint32_t rb, bool mirrored) { s.put(f, *s.ex.A, ra, *s.ex.B, rb, mirrored ? ct : cd); });
  });
  s.host_s += secs_since(th);
  return s.c->stop();
}
int relate_points_cb(void* u, const uint32_t* pairs, const uint8_t* where, size_t n) {
  RelateSink& s = *(RelateSink*)u;
  auto th = Clock::now();
  const Layer* P = s.swapped ? s.ex.B : s.ex.A;
  const Layer* G = s.swapped ? s.ex.A : s.ex.B;
  s.total += s.c->send_rows(n, [&](size_t i, RowChunk& f) {
    const uint32_t code = where[i] == 2 ? GSJ_DE9IM_PT_BOUNDARY : GSJ_DE9IM_PT_INTERIOR;   // relate(point, polygon)
    G->members(pairs[2 * i + 1], [&](uint32_t rg) {
      if (s.swapped) s.put(f, *G, rg, *P, pairs[2 * i], gsj_de9im_transpose(code));
      else s.put(f, *P, pairs[2 * i], *G, rg, code);
    });
  });
  s.host_s += secs_since(th);
  return s.c->stop();
}
void op_relate(Reply& c, const std::vector<std::string>& a) {
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
  c.open_stream();
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
  c.close_stream(stats);
  logf("relate %s x %s: %s%s", a[0].c_str(), a[1].c_str(), stats.c_str(), c.peer().gone ? " (client went away)" : "");
}

// DWITHIN (reference op_dwithin, daemon/gpujoind.cu:396-477)
struct DistSink {
  Reply* c;
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
      s.total += s.c->send_rows(n, [&](size_t i, RowChunk& f) {
        const double v = (s.with_distance && dist) ? dist[i] : NaN;
        G->members(pairs[2 * i + 1], [&](uint32_t rg) {
          const uint32_t rp = pairs[2 * i];
          if (s.swapped) { f.id(G->ids.data(rg), G->ids.len(rg)); f.id(P->ids.data(rp), P->ids.len(rp)); }
          else { f.id(P->ids.data(rp), P->ids.len(rp)); f.id(G->ids.data(rg), G->ids.len(rg)); }
          f.f64(v);
          f.end_row();
        });
      });
    }
  } else if (s.stats_only) {
    uint64_t t = 0;
    for (size_t i = 0; i < n; i++) t += s.ex.count(pairs[2 * i], pairs[2 * i + 1]);
    s.total += t;
  } else {
    s.total += s.c->send_rows(n, [&](size_t i, RowChunk& f) {
      const double v = (s.with_distance && dist) ? dist[i] : NaN;
      s.ex.rows(pairs[2 * i], pairs[2 * i + 1], [&](uint32_t ra, uint32_t rb, bool) {
        f.id(s.ex.A->ids.data(ra), s.ex.A->ids.len(ra)); f.id(s.ex.B->ids.data(rb), s.ex.B->ids.len(rb)); f.f64(v); f.end_row();
      });
    });
  }
  s.host_s += secs_since(th);
  return s.c->stop();
}
void op_dwithin(Reply& c, const std::vector<std::string>& a) {
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
  c.open_stream();
  auto t0 = Clock::now();
  DistSink s{&c, Expand{lp.A.get(), lp.B.get(), lp.same, distinct}, stats_only, points, lp.swapped, with_distance && !erode};
  gsj_request r{};
  r.dist_cb = dist_rows_cb;
  r.user = &s;
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
        r.need_listed = 1;
        r.need_list = lp.A->self_needed.data();
        r.n_need = lp.A->self_needed.size();
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
  c.close_stream(stats);
  logf("dwithin %s x %s %s %s %s: %s%s", a[0].c_str(), a[1].c_str(), mode.c_str(), a[3].c_str(), a[4].c_str(), stats.c_str(),
       c.peer().gone ? " (client went away)" : "");
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

// ------------------------------------------------------------------------------------------------- dispatch
using OpFn = void (*)(Reply&, const std::vector<std::string>&);

void op_load(Reply& r, const std::vector<std::string>& a) {
  if (a.size() < 7) throw std::runtime_error("LOAD needs layer, source, id_col, geom_col, filter, subdivide, dedup");
  r.json(build_layer(a[0], a[1], to_int(a[5], 0), to_bool(a[6], true)));
}
void op_load_query(Reply& r, const std::vector<std::string>& a) {
  if (a.size() < 4) throw std::runtime_error("LOAD_QUERY needs layer, source, subdivide, dedup");
  r.json(build_layer(a[0], a[1], to_int(a[2], 0), to_bool(a[3], true)));
}
void op_drop(Reply& r, const std::vector<std::string>& a) {
  if (a.empty()) throw std::runtime_error("DROP needs layer");
  std::shared_ptr<Layer> old;   // released after the registry lock; running requests keep the device layer
  {
    std::lock_guard<std::mutex> g(g_reg);
    auto it = g_layers.find(a[0]);
    if (it != g_layers.end()) { old = it->second; g_layers.erase(it); }
  }
  r.json(old ? "true" : "false");
}

const std::unordered_map<uint32_t, OpFn>& op_table() {
  static const std::unordered_map<uint32_t, OpFn> table = {
      {GSP_OP_LOAD, op_load},
      {GSP_OP_LOAD_QUERY, op_load_query},
      {GSP_OP_DROP, op_drop},
      {GSP_OP_LIST, [](Reply& r, const std::vector<std::string>&) { r.json(op_list()); }},
      {GSP_OP_STATUS, [](Reply& r, const std::vector<std::string>&) { r.json(op_status()); }},
      {GSP_OP_JOIN, op_join},
      {GSP_OP_CLUSTER, op_cluster},
      {GSP_OP_RELATE, op_relate},
      {GSP_OP_DWITHIN, op_dwithin},
  };
  return table;
}

std::atomic<int> g_connections{0};
constexpr int MAX_CONNECTIONS = 2048;

// one connection: one request, one response (the extension opens a connection per call)
void serve_connection(int fd) {
  Peer peer(fd);
  struct Count {   // connections in flight, for the cap in accept_loop
    ~Count() { g_connections--; }
  } count;
  WireRequest q;
  if (!read_request(peer, q)) return;   // malformed, oversized or truncated: dropped without a reply
  Reply rep(peer);
  if (!token_matches(q.token)) {
    logf("rejected request: bad token");
    rep.fail("authentication failed");
    return;
  }
  try {
    const auto& ops = op_table();
    const auto it = ops.find(q.op);
    if (it == ops.end()) throw std::runtime_error("unknown op " + std::to_string(q.op));
    it->second(rep, q.args);
  } catch (const std::exception& e) {
    logf("error: %s", e.what());
    rep.fail(e.what());
  }
}

// ------------------------------------------------------------------------------------------------- listening
class Listener {
 public:
  void unix_socket(const std::string& path, mode_t mode) {
    sockaddr_un addr{};
    addr.sun_family = AF_UNIX;
    if (path.size() >= sizeof addr.sun_path) throw std::runtime_error("unix socket path too long: " + path);
    memcpy(addr.sun_path, path.c_str(), path.size() + 1);
    ::unlink(path.c_str());
    add(AF_UNIX, reinterpret_cast<sockaddr*>(&addr), sizeof addr, "unix:" + path);
    if (::chmod(path.c_str(), mode) != 0) throw std::runtime_error("chmod " + path + ": " + strerror(errno));
  }
  void tcp(const std::string& host_port) {
    const size_t colon = host_port.rfind(':');
    if (colon == std::string::npos) throw std::runtime_error("tcp address must be HOST:PORT");
    sockaddr_in addr{};
    addr.sin_family = AF_INET;
    addr.sin_port = htons((uint16_t)std::stoi(host_port.substr(colon + 1)));
    if (inet_pton(AF_INET, host_port.substr(0, colon).c_str(), &addr.sin_addr) != 1) throw std::runtime_error("bad tcp address " + host_port);
    add(AF_INET, reinterpret_cast<sockaddr*>(&addr), sizeof addr, "tcp:" + host_port);
  }
  // accept forever; every connection gets a thread of its own, up to MAX_CONNECTIONS at once
  [[noreturn]] void accept_loop() {
    for (;;) {
      if (::poll(fds_.data(), fds_.size(), -1) < 0) {
        if (errno == EINTR) continue;
        throw std::runtime_error(std::string("poll: ") + strerror(errno));
      }
      for (pollfd& l : fds_) {
        if (!(l.revents & POLLIN)) continue;
        const int fd = ::accept(l.fd, nullptr, nullptr);
        if (fd < 0) continue;
        const int on = 1;
        setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &on, sizeof on);   // no effect on unix sockets
        if (g_connections.fetch_add(1) >= MAX_CONNECTIONS) {
          g_connections--;
          Peer busy(fd);
          busy.write_text(GSP_ERR, "daemon busy: too many connections");
          continue;
        }
        std::thread(serve_connection, fd).detach();
      }
    }
  }

 private:
  std::vector<pollfd> fds_;
  void add(int family, const sockaddr* addr, socklen_t len, const std::string& what) {
    const int fd = ::socket(family, SOCK_STREAM, 0);
    if (fd < 0) throw std::runtime_error(what + ": " + strerror(errno));
    const int on = 1;
    if (family == AF_INET) setsockopt(fd, SOL_SOCKET, SO_REUSEADDR, &on, sizeof on);
    if (::bind(fd, addr, len) != 0 || ::listen(fd, 128) != 0) {
      const std::string err = strerror(errno);
      ::close(fd);
      throw std::runtime_error(what + ": " + err);
    }
    fds_.push_back({fd, POLLIN, 0});
    logf("listening on %s", what.c_str());
  }
};

}  // namespace

// ------------------------------------------------------------------------------------------------- main
int main(int argc, char** argv) {
  struct Options {
    std::string socket = "/tmp/gpujoind.sock", tcp, token_file;
    int lanes = 4, device = 0, batch = 16;
    uint64_t scratch_mb = 0;
    mode_t socket_mode = 0660;
  } o;
  const std::vector<std::pair<std::string, std::function<void(const char*)>>> flags = {
      {"--socket", [&](const char* v) { o.socket = v; }},
      {"--socket-mode", [&](const char* v) { o.socket_mode = (mode_t)std::stoul(v, nullptr, 8); }},
      {"--tcp", [&](const char* v) { o.tcp = v; }},
      {"--token-file", [&](const char* v) { o.token_file = v; }},
      {"--lanes", [&](const char* v) { o.lanes = std::stoi(v); }},
      {"--threads", [&](const char* v) { g_threads = std::stoi(v); }},
      {"--data-dir", [&](const char* v) { g_data_dir = v; }},
      {"--device", [&](const char* v) { o.device = std::stoi(v); }},
      {"--scratch-mb", [&](const char* v) { o.scratch_mb = std::stoull(v); }},
      {"--max-rows", [&](const char* v) { g_max_rows = std::stoull(v); }},
      {"--batch", [&](const char* v) { o.batch = std::stoi(v); }},
  };
  auto usage = [&] {
    std::string u = "usage: gpujoind";
    for (const auto& f : flags) u += " [" + f.first + " V]";
    fprintf(stderr, "%s\n", u.c_str());
    return 1;
  };
  for (int i = 1; i < argc; i++) {
    const auto f = std::find_if(flags.begin(), flags.end(), [&](const auto& x) { return x.first == argv[i]; });
    if (f == flags.end() || i + 1 >= argc) return usage();
    try {
      f->second(argv[++i]);
    } catch (const std::exception&) {
      return usage();
    }
  }
  if (!o.token_file.empty()) {
    std::ifstream tf(o.token_file);
    std::getline(tf, g_token);
    if (g_token.empty()) { fprintf(stderr, "gpujoind: empty token file %s\n", o.token_file.c_str()); return 1; }
  }
  if (!o.tcp.empty() && g_token.empty()) { fprintf(stderr, "gpujoind: --tcp requires --token-file\n"); return 1; }
  char real[PATH_MAX];
  if (!::realpath(g_data_dir.c_str(), real)) { fprintf(stderr, "gpujoind: data dir %s: %s\n", g_data_dir.c_str(), strerror(errno)); return 1; }
  g_data_real = real;
  signal(SIGPIPE, SIG_IGN);
  try {
    check(gsj_ctx_create(o.device, &g_ctx), "device");
    check(gsj_sched_create(g_ctx, o.lanes, o.scratch_mb << 20, &g_sched), "scheduler");
    check(gsj_sched_set_batch(g_sched, o.batch), "batching");
    logf("gpujoind on %s, %d lanes, data in %s", gsj_ctx_status_json(g_ctx), o.lanes, g_data_real.c_str());
    Listener l;
    l.unix_socket(o.socket, o.socket_mode);
    if (!o.tcp.empty()) l.tcp(o.tcp);
    l.accept_loop();
  } catch (const std::exception& e) {
    fprintf(stderr, "gpujoind: %s\n", e.what());
    return 1;
  }
}
