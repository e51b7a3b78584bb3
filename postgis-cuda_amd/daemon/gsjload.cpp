// gsjload: open-loop load client for gpujoind (the role of the reference's scripts/bench_amortized.py, in C++ so that
// draining row streams does not limit what is measured). It reads a request list, sends each request at its arrival
// time on a connection of its own (as the extension does, one request per connection), reads the whole response, and
// reports per request the latency from the last request byte to the end-of-stream stats frame (or the JSON reply),
// the rows received and the daemon's stats, then the p50 / p99 latency and the wall time of the run.
//
//   gsjload --socket PATH [--token T] < requests
//
// Request lines: `t op arg1 arg2 ...`, fields separated by tabs: t = arrival time in seconds from the start, op = the
// protocol's op number. Output: one JSON object per request, then one summary object.
#include "protocol.h"

#include <sys/socket.h>
#include <sys/un.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  double t = 0;
  uint32_t op = 0;
  std::vector<std::string> args;
};

struct Result {
  bool ok = false;
  double latency_s = 0, start_s = 0;
  uint64_t rows = 0, bytes = 0;
  std::string stats, error;
};

class Stream {   // buffered reads from a socket
 public:
  explicit Stream(int fd) : fd_(fd), buf_(1 << 20) {}
  bool read(void* dst, size_t n) {
    auto* p = static_cast<char*>(dst);
    while (n) {
      if (pos_ == len_) {
        const ssize_t k = ::recv(fd_, buf_.data(), buf_.size(), 0);
        if (k <= 0) return false;
        len_ = (size_t)k;
        pos_ = 0;
        bytes += (uint64_t)k;
      }
      const size_t take = std::min(n, len_ - pos_);
      memcpy(p, buf_.data() + pos_, take);
      pos_ += take; p += take; n -= take;
    }
    return true;
  }
  bool skip(size_t n) {
    char tmp[256];
    while (n) {
      const size_t k = std::min(n, sizeof tmp);
      if (!read(tmp, k)) return false;
      n -= k;
    }
    return true;
  }
  bool u32(uint32_t& v) { return read(&v, 4); }
  // a whole row already in the buffer ((u16 len, id) x ids, then `tail` bytes): step over it in place; false (nothing
  // consumed) when it runs past the buffered bytes
  bool skip_row_buffered(int ids, size_t tail) {
    size_t p = pos_;
    for (int k = 0; k < ids; k++) {
      if (len_ - p < 2) return false;
      uint16_t l;
      memcpy(&l, buf_.data() + p, 2);
      p += 2;
      if (len_ - p < l) return false;
      p += l;
    }
    if (len_ - p < tail) return false;
    pos_ = p + tail;
    return true;
  }
  bool text(std::string& s) {
    uint32_t l;
    if (!u32(l)) return false;
    s.resize(l);
    return l == 0 || read(&s[0], l);
  }
  uint64_t bytes = 0;

 private:
  int fd_;
  std::vector<char> buf_;
  size_t pos_ = 0, len_ = 0;
};

// rows of one frame: (u16 len, id) pairs and the op's fixed tail
bool skip_row(Stream& s, uint32_t op) {
  const int ids = op == GSP_OP_CLUSTER ? 1 : 2;
  if (s.skip_row_buffered(ids, op == GSP_OP_CLUSTER ? 4 : (op == GSP_OP_RELATE ? 9 : 8))) return true;
  for (int k = 0; k < ids; k++) {
    uint16_t l;
    if (!s.read(&l, 2) || !s.skip(l)) return false;
  }
  return s.skip(op == GSP_OP_CLUSTER ? 4 : (op == GSP_OP_RELATE ? 9 : 8));
}

Result run_one(const std::string& sock, const std::string& token, const Request& q) {
  Result r;
  const int fd = ::socket(AF_UNIX, SOCK_STREAM, 0);
  sockaddr_un a{};
  a.sun_family = AF_UNIX;
  strncpy(a.sun_path, sock.c_str(), sizeof a.sun_path - 1);
  if (fd < 0 || ::connect(fd, reinterpret_cast<sockaddr*>(&a), sizeof a) != 0) {
    r.error = std::string("connect: ") + strerror(errno);
    if (fd >= 0) ::close(fd);
    return r;
  }
  std::string msg;
  auto put = [&](uint32_t v) { msg.append(reinterpret_cast<const char*>(&v), 4); };
  put(GSP_MAGIC);
  put(q.op);
  put((uint32_t)q.args.size() + 1);
  put((uint32_t)token.size());
  msg += token;
  for (const std::string& x : q.args) { put((uint32_t)x.size()); msg += x; }
  for (size_t off = 0; off < msg.size();) {
    const ssize_t k = ::send(fd, msg.data() + off, msg.size() - off, MSG_NOSIGNAL);
    if (k <= 0) { r.error = "send failed"; ::close(fd); return r; }
    off += (size_t)k;
  }
  const Clock::time_point t0 = Clock::now();   // the last request byte is out
  Stream s(fd);
  uint32_t status = 0;
  if (!s.u32(status)) r.error = "connection closed";
  else if (status == GSP_ERR) { s.text(r.error); if (r.error.empty()) r.error = "error"; }
  else if (q.op == GSP_OP_LOAD || q.op == GSP_OP_LOAD_QUERY || q.op == GSP_OP_DROP || q.op == GSP_OP_LIST || q.op == GSP_OP_STATUS) {
    r.ok = s.text(r.stats);
  } else {
    for (;;) {
      uint32_t n;
      if (!s.u32(n)) { r.error = "stream cut"; break; }
      if (n == GSP_FRAME_ERR) { s.text(r.error); break; }
      if (n == 0) { r.ok = s.text(r.stats); break; }
      for (uint32_t i = 0; i < n; i++)
        if (!skip_row(s, q.op)) { r.error = "stream cut"; break; }
      if (!r.error.empty()) break;
      r.rows += n;
    }
  }
  r.latency_s = std::chrono::duration<double>(Clock::now() - t0).count();
  r.bytes = s.bytes;
  ::close(fd);
  return r;
}

std::string jstr(const std::string& s) {
  std::string o = "\"";
  for (unsigned char c : s) {
    if (c == '"' || c == '\\') { o += '\\'; o += (char)c; }
    else if (c < 0x20) { char b[8]; snprintf(b, sizeof b, "\\u%04x", c); o += b; }
    else o += (char)c;
  }
  return o + "\"";
}

}  // namespace

int main(int argc, char** argv) {
  std::string sock, token;
  for (int i = 1; i + 1 < argc; i += 2) {
    const std::string k = argv[i];
    if (k == "--socket") sock = argv[i + 1];
    else if (k == "--token") token = argv[i + 1];
    else { fprintf(stderr, "usage: gsjload --socket PATH [--token T] < requests\n"); return 1; }
  }
  if (sock.empty()) { fprintf(stderr, "usage: gsjload --socket PATH [--token T] < requests\n"); return 1; }
  std::vector<Request> reqs;
  for (std::string line; std::getline(std::cin, line);) {
    if (line.empty()) continue;
    std::vector<std::string> f;
    std::stringstream ss(line);
    for (std::string x; std::getline(ss, x, '\t');) f.push_back(x);
    if (f.size() < 2) { fprintf(stderr, "bad request line: %s\n", line.c_str()); return 1; }
    Request q;
    q.t = std::stod(f[0]);
    q.op = (uint32_t)std::stoul(f[1]);
    q.args.assign(f.begin() + 2, f.end());
    reqs.push_back(std::move(q));
  }
  std::vector<Result> res(reqs.size());
  std::vector<std::thread> th;
  th.reserve(reqs.size());
  const Clock::time_point start = Clock::now();
  for (size_t i = 0; i < reqs.size(); i++) {
    std::this_thread::sleep_until(start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(reqs[i].t)));
    th.emplace_back([&, i] {
      res[i] = run_one(sock, token, reqs[i]);
      res[i].start_s = reqs[i].t;
    });
  }
  for (auto& t : th) t.join();
  const double wall = std::chrono::duration<double>(Clock::now() - start).count();
  std::vector<double> lat;
  uint64_t rows = 0, failed = 0;
  for (size_t i = 0; i < res.size(); i++) {
    const Result& r = res[i];
    printf("{\"i\":%zu,\"ok\":%s,\"latency_ms\":%.4f,\"rows\":%llu,\"bytes\":%llu,\"stats\":%s%s%s}\n", i, r.ok ? "true" : "false", r.latency_s * 1e3,
           (unsigned long long)r.rows, (unsigned long long)r.bytes, r.ok && !r.stats.empty() ? r.stats.c_str() : "null",
           r.error.empty() ? "" : ",\"error\":", r.error.empty() ? "" : jstr(r.error).c_str());
    if (r.ok) { lat.push_back(r.latency_s * 1e3); rows += r.rows; } else failed++;
  }
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double p) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(p / 100.0 * (double)lat.size()))]; };
  printf("{\"summary\":true,\"requests\":%zu,\"failed\":%llu,\"wall_s\":%.4f,\"rows\":%llu,\"p50_ms\":%.4f,\"p99_ms\":%.4f}\n", res.size(),
         (unsigned long long)failed, wall, (unsigned long long)rows, pct(50), pct(99));
  return failed ? 2 : 0;
}
