/* Test infrastructure: a replay of how the PostgreSQL extension reads gpujoind streams (ext/gpu_spatial.c:120-470),
 * written here from that reader's behaviour so that this daemon's streams are checked by a reader with the
 * extension's constraints rather than by this build's own Python client:
 *   - one connection per request; request = magic, op, nargs + 1, token, args (u32 length + bytes each);
 *   - a fixed 1 MiB receive buffer, refilled by poll + read after moving the unread tail to its front; a single take
 *     may not exceed the buffer, strings are copied out in pieces of half of it;
 *   - status u32 (GSP_ERR: message); then frames of u32 nrows rows, ended by 0, or GSP_FRAME_ERR + message; then
 *     the stats JSON;
 *   - rows by kind: join (u16 id, u16 id, f64 area, NaN = NULL), relate (u16 id, u16 id, char[9]), cluster (u16 id,
 *     u32 cluster read as int32), pair (u16 id, u16 id, 8 bytes skipped: DWITHIN); "stats": no row may come.
 * Output: one line per row, fields separated by tabs (area as %.17g or NULL), then "#stats <json>"; an error
 * prints "ERROR: gpujoind: <message>" (as the extension raises it) and exits 1.
 * usage: srf_replay SOCKET TOKEN KIND OP ARGS... */
#include "../postgis-cuda_amd/daemon/protocol.h"

#include <errno.h>
#include <math.h>
#include <poll.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/socket.h>
#include <sys/un.h>
#include <unistd.h>

typedef struct {
  int fd;
  size_t pos, len;
  char buf[1 << 20];
} Conn;

static void fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fputs("ERROR: ", stdout);
  vprintf(fmt, ap);
  fputc('\n', stdout);
  va_end(ap);
  exit(1);
}

static void refill(Conn* c) {
  if (c->pos > 0 && c->pos < c->len) memmove(c->buf, c->buf + c->pos, c->len - c->pos);
  c->len -= c->pos;
  c->pos = 0;
  for (;;) {
    struct pollfd p = {c->fd, POLLIN, 0};
    const int r = poll(&p, 1, 100);
    if (r < 0 && errno != EINTR) fail("gpujoind: poll failed: %s", strerror(errno));
    if (r <= 0) continue;
    const ssize_t k = read(c->fd, c->buf + c->len, sizeof c->buf - c->len);
    if (k == 0) fail("gpujoind closed the connection unexpectedly");
    if (k < 0) {
      if (errno == EINTR || errno == EAGAIN) continue;
      fail("gpujoind: read failed: %s", strerror(errno));
    }
    c->len += (size_t)k;
    return;
  }
}

static const char* take(Conn* c, size_t n) {
  if (n > sizeof c->buf) fail("take of %zu bytes exceeds the receive buffer", n);
  while (c->len - c->pos < n) refill(c);
  const char* p = c->buf + c->pos;
  c->pos += n;
  return p;
}

static uint32_t u32(Conn* c) {
  uint32_t v;
  memcpy(&v, take(c, 4), 4);
  return v;
}

static char* str(Conn* c) {
  const uint32_t n = u32(c);
  char* s = malloc((size_t)n + 1);
  uint32_t got = 0;
  while (got < n) {
    const uint32_t k = n - got < (uint32_t)(sizeof c->buf / 2) ? n - got : (uint32_t)(sizeof c->buf / 2);
    memcpy(s + got, take(c, k), k);
    got += k;
  }
  s[n] = 0;
  return s;
}

static void put_u32(char** p, uint32_t v) {
  memcpy(*p, &v, 4);
  *p += 4;
}

static void send_all(int fd, const char* p, size_t n) {
  while (n) {
    const ssize_t k = send(fd, p, n, MSG_NOSIGNAL);
    if (k < 0) {
      if (errno == EINTR) continue;
      fail("gpujoind: send failed: %s", strerror(errno));
    }
    p += k;
    n -= (size_t)k;
  }
}

static void print_id(Conn* c) {
  uint16_t l;
  memcpy(&l, take(c, 2), 2);
  fwrite(take(c, l), 1, l, stdout);
}

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: srf_replay SOCKET TOKEN KIND OP ARGS...\n");
    return 2;
  }
  const char* kind = argv[3];
  const uint32_t op = (uint32_t)atoi(argv[4]);
  const int nargs = argc - 5;
  static Conn conn;
  Conn* c = &conn;
  c->fd = socket(AF_UNIX, SOCK_STREAM, 0);
  struct sockaddr_un a;
  memset(&a, 0, sizeof a);
  a.sun_family = AF_UNIX;
  strncpy(a.sun_path, argv[1], sizeof a.sun_path - 1);
  if (c->fd < 0 || connect(c->fd, (struct sockaddr*)&a, sizeof a) != 0) fail("could not connect to gpujoind at \"%s\"", argv[1]);
  size_t total = 16 + strlen(argv[2]);
  for (int i = 0; i < nargs; i++) total += 4 + strlen(argv[5 + i]);
  char* req = malloc(total);
  char* p = req;
  put_u32(&p, GSP_MAGIC);
  put_u32(&p, op);
  put_u32(&p, (uint32_t)nargs + 1);
  put_u32(&p, (uint32_t)strlen(argv[2]));
  memcpy(p, argv[2], strlen(argv[2]));
  p += strlen(argv[2]);
  for (int i = 0; i < nargs; i++) {
    const size_t l = strlen(argv[5 + i]);
    put_u32(&p, (uint32_t)l);
    memcpy(p, argv[5 + i], l);
    p += l;
  }
  send_all(c->fd, req, total);
  free(req);
  if (u32(c) != GSP_OK) fail("gpujoind: %s", str(c));
  for (;;) {
    const uint32_t n = u32(c);
    if (n == 0) break;
    if (n == GSP_FRAME_ERR) fail("gpujoind: %s", str(c));
    if (!strcmp(kind, "stats")) fail("gpujoind: unexpected rows in stats mode");
    for (uint32_t i = 0; i < n; i++) {
      if (!strcmp(kind, "cluster")) {
        print_id(c);
        uint32_t k;
        memcpy(&k, take(c, 4), 4);
        printf("\t%d\n", (int32_t)k);
        continue;
      }
      print_id(c);
      putchar('\t');
      print_id(c);
      if (!strcmp(kind, "join")) {
        double v;
        memcpy(&v, take(c, 8), 8);
        if (isnan(v)) printf("\tNULL\n");
        else printf("\t%.17g\n", v);
      } else if (!strcmp(kind, "relate")) {
        printf("\t%.9s\n", take(c, 9));
      } else {
        (void)take(c, 8);
        putchar('\n');
      }
    }
  }
  printf("#stats %s\n", str(c));
  close(c->fd);
  return 0;
}
