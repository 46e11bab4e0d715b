// Trajectory output in the reference's xyz text format, staged
// asynchronously: io_util::WriteXYZForm (io_util.hpp:25-31) writes a count
// line, "time T" and one FPDPD::writeAscii line per bead (ptcl_class.hpp:
// 67-73: type letter, pos, id, prop, amp_id, unit, vel, vel_buf, acc, each
// real as %.15g), in upload order.
//
// mdpd_write_xyz_async gathers the columns on the device into upload order,
// copies them into pinned host memory on the context's stream and returns;
// a host thread waits for that copy, formats the lines in parallel chunks
// and writes the file, while the caller goes on stepping.  The next write,
// mdpd_io_wait and mdpd_destroy wait for it.  The writer thread makes no HIP
// call: a host function enqueued behind the copy signals it, so it cannot
// disturb a capture of the context's stream (mdpd_run records its steps as
// graphs on that stream) that the caller starts meanwhile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"

namespace mdpd {

struct XyzWriter {
  XyzRec* dev = nullptr;   // device staging, upload order
  XyzRec* host = nullptr;  // pinned copy
  uint32_t cap = 0;        // records both buffers hold
  std::mutex mu;
  std::condition_variable cv;
  bool copied = false;     // the pinned copy of the current write is complete
  std::thread thr;
  int status = MDPD_OK;    // of the last finished write
  std::string err;
};

namespace {

// Parameter::atom_type (parameter.hpp:207-209)
const char kAtomType[] = "ONCSPZXONCSPZXONCSPZX";

void format_rows(const XyzRec* r, uint32_t b, uint32_t e, std::string& out) {
  char line[512];
  out.reserve((size_t)(e - b) * 300);
  for (uint32_t i = b; i < e; ++i) {
    const XyzRec& q = r[i];
    const char t = q.prop < sizeof(kAtomType) - 1 ? kAtomType[q.prop] : '?';
    const int k = std::snprintf(
        line, sizeof(line),
        "%c %.15g %.15g %.15g %u %u %u %u %.15g %.15g %.15g %.15g %.15g %.15g %.15g %.15g %.15g\n", t,
        q.pos[0], q.pos[1], q.pos[2], q.id, q.prop, q.amp_id, q.unit, q.vel[0], q.vel[1], q.vel[2],
        q.vbuf[0], q.vbuf[1], q.vbuf[2], (double)q.acc[0], (double)q.acc[1], (double)q.acc[2]);
    out.append(line, (size_t)std::min(k, (int)sizeof(line) - 1));
  }
}

// enqueued behind the copy on the context's stream (no HIP call in here)
void xyz_copied(void* p) {
  XyzWriter* w = static_cast<XyzWriter*>(p);
  {
    std::lock_guard<std::mutex> lk(w->mu);
    w->copied = true;
  }
  w->cv.notify_all();
}

// the writer thread: wait for the copy, format, write
void write_job(XyzWriter* w, std::string path, uint32_t n, uint32_t time, bool append) {
  {
    std::unique_lock<std::mutex> lk(w->mu);
    w->cv.wait(lk, [w] { return w->copied; });
  }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const uint32_t nt = std::max(1u, std::min({16u, hw, (n + 16383u) / 16384u}));
  std::vector<std::string> part(nt);
  std::vector<std::thread> pool;
  for (uint32_t k = 0; k < nt; ++k) {
    const uint32_t b = (uint32_t)((uint64_t)n * k / nt), e = (uint32_t)((uint64_t)n * (k + 1) / nt);
    pool.emplace_back(format_rows, w->host, b, e, std::ref(part[k]));
  }
  for (auto& t : pool) t.join();
  FILE* f = std::fopen(path.c_str(), append ? "a" : "w");
  if (!f) {
    w->status = MDPD_ERR_ARG;
    w->err = "cannot open " + path;
    return;
  }
  bool ok = std::fprintf(f, "%u\ntime %u\n", n, time) > 0;
  for (const std::string& s : part) ok = ok && std::fwrite(s.data(), 1, s.size(), f) == s.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    w->status = MDPD_ERR_ARG;
    w->err = "write to " + path + " failed";
  }
}

int join(mdpd_ctx* c) {
  XyzWriter* w = c->xyz;
  if (!w) return MDPD_OK;
  if (w->thr.joinable()) w->thr.join();
  const int rc = w->status;
  w->status = MDPD_OK;
  if (rc) return fail(c, rc, w->err);
  return MDPD_OK;
}

}  // namespace

void xyz_writer_free(XyzWriter* w) {
  if (!w) return;
  if (w->thr.joinable()) w->thr.join();
  if (w->dev) (void)hipFree(w->dev);
  if (w->host) (void)hipHostFree(w->host);
  delete w;
}

}  // namespace mdpd

using namespace mdpd;

extern "C" {

int mdpd_write_xyz_async(mdpd_ctx* c, const char* path, uint32_t time, int32_t append) {
  if (!c || !path) return fail(c, MDPD_ERR_ARG, "null argument");
  int rc = join(c);  // the previous write (and its error, if it failed)
  if (rc) return rc;
  if (!c->xyz) c->xyz = new XyzWriter();
  XyzWriter* w = c->xyz;
  const DevState& d = c->d;
  const uint32_t n = d.n;
  if (n > w->cap) {
    if (w->dev) (void)hipFree(w->dev);
    if (w->host) (void)hipHostFree(w->host);
    w->dev = nullptr;
    w->host = nullptr;
    w->cap = 0;
    HIPCHK(c, hipMalloc(&w->dev, sizeof(XyzRec) * (size_t)n));
    HIPCHK(c, hipHostMalloc(&w->host, sizeof(XyzRec) * (size_t)n, hipHostMallocDefault));
    w->cap = n;
  }
  // after a stand-alone drift the predicted velocity is the reference's vel
  launch_xyz_gather(d, c->drifted ? d.vbuf : d.vel, w->dev, c->st);
  HIPCHK(c, hipGetLastError());
  if (n) HIPCHK(c, hipMemcpyAsync(w->host, w->dev, sizeof(XyzRec) * (size_t)n, hipMemcpyDeviceToHost, c->st));
  {
    std::lock_guard<std::mutex> lk(w->mu);
    w->copied = false;
  }
  HIPCHK(c, hipLaunchHostFunc(c->st, xyz_copied, w));
  w->status = MDPD_OK;
  w->thr = std::thread(write_job, w, std::string(path), n, time, append != 0);
  return MDPD_OK;
}

int mdpd_io_wait(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  return join(c);
}

// ---------------------------------------------------------------- reading
// io_util::ReadXYZForm (io_util.hpp:103-135) with FPDPD::readAscii
// (ptcl_class.hpp:60-65): "%u" count line, "time %u" line, then one line of
// "%c %lf %lf %lf %u %u %u %u %lf x9" per bead; the other fields of the
// record (delta_sumr, press, nei_cm_pos, density) are zero.  The file is
// read once; line starts are found and lines parsed in parallel chunks with
// strtod / strtoul (the conversions of fscanf's %lf / %u).

static int xyz_header(const char* b, const char* e, uint32_t* n, uint32_t* time, const char** body,
                      std::string& err) {
  // "%u\n" then "%s %u\n"
  char* q = nullptr;
  const char* p = b;
  while (p < e && std::isspace((unsigned char)*p)) ++p;
  const unsigned long nv = std::strtoul(p, &q, 10);
  if (q == p) {
    err = "cannot read the number of particles";
    return MDPD_ERR_PARSE;
  }
  p = q;
  while (p < e && std::isspace((unsigned char)*p)) ++p;
  const char* tag = p;
  while (p < e && !std::isspace((unsigned char)*p)) ++p;
  if (std::string(tag, p) != "time") {
    err = "The file format of init_config.xyz may be old.";
    return MDPD_ERR_PARSE;
  }
  const unsigned long tv = std::strtoul(p, &q, 10);
  if (q == p || nv > 0xffffffffUL || tv > 0xffffffffUL) {
    err = "The file format of init_config.xyz may be old.";
    return MDPD_ERR_PARSE;
  }
  p = q;
  while (p < e && *p != '\n') ++p;
  if (p < e) ++p;
  *n = (uint32_t)nv;
  *time = (uint32_t)tv;
  *body = p;
  return MDPD_OK;
}

// one record line [p, e); false if it does not hold the 17 columns
static bool xyz_line(const char* p, const char* e, char* rec) {
  while (p < e && (*p == ' ' || *p == '\t')) ++p;
  if (p >= e) return false;
  ++p;  // the type letter (%c)
  double v[12];
  uint32_t u[4];
  char* q = nullptr;
  for (int k = 0; k < 3; ++k) {
    v[k] = std::strtod(p, &q);
    if (q == p || q > e) return false;  // a short line must not run into the next
    p = q;
  }
  for (int k = 0; k < 4; ++k) {
    const unsigned long x = std::strtoul(p, &q, 10);
    if (q == p || q > e) return false;
    u[k] = (uint32_t)x;
    p = q;
  }
  for (int k = 3; k < 12; ++k) {
    v[k] = std::strtod(p, &q);
    if (q == p || q > e) return false;
    p = q;
  }
  std::memset(rec, 0, MDPD_FPDPD_STRIDE);
  std::memcpy(rec + 0, u, 16);          // id, prop, amp_id, unit
  std::memcpy(rec + 16, v, 24);         // pos
  std::memcpy(rec + 64, v + 3, 72);     // vel, vel_buf, acc
  return true;
}

int mdpd_read_xyz_header(const char* path, uint32_t* n, uint32_t* time) {
  if (!path || !n || !time) return fail(nullptr, MDPD_ERR_ARG, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(nullptr, MDPD_ERR_PARSE, std::string("cannot open ") + path);
  // one spare byte: the parse below (strtoul) stops at a NUL at the latest
  char buf[257];
  const size_t k = std::fread(buf, 1, sizeof(buf) - 1, f);
  std::fclose(f);
  buf[k] = '\0';
  const char* body = nullptr;
  std::string err;
  const int rc = xyz_header(buf, buf + k, n, time, &body, err);
  if (rc) return fail(nullptr, rc, err);
  return MDPD_OK;
}

int mdpd_read_xyz(const char* path, void* aos, uint32_t cap, size_t stride, uint32_t* n_out,
                  uint32_t* time_out) {
  if (!path || (!aos && cap) || !n_out || !time_out) return fail(nullptr, MDPD_ERR_ARG, "null argument");
  if (stride < MDPD_FPDPD_STRIDE) return fail(nullptr, MDPD_ERR_ARG, "stride < 232");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(nullptr, MDPD_ERR_PARSE, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<char> data((size_t)std::max(sz, 0L) + 1, '\0');
  const size_t got = std::fread(data.data(), 1, (size_t)std::max(sz, 0L), f);
  std::fclose(f);
  const char* b = data.data();
  const char* e = b + got;
  uint32_t n = 0, time = 0;
  const char* body = nullptr;
  std::string err;
  int rc = xyz_header(b, e, &n, &time, &body, err);
  if (rc) return fail(nullptr, rc, err);
  if (n > cap) return fail(nullptr, MDPD_ERR_CAPACITY,
                           "# of particles specified in init_config.xyz exceeds maximum buffer.");
  // line starts, in parallel chunks (blank lines are not records)
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t len = (size_t)(e - body);
  const unsigned nt = (unsigned)std::max<size_t>(1, std::min<size_t>({16u, hw, len / (1u << 20) + 1}));
  std::vector<std::vector<const char*>> starts(nt);
  {
    std::vector<std::thread> pool;
    for (unsigned k = 0; k < nt; ++k)
      pool.emplace_back([&, k] {
        const char* lo = body + len * k / nt;
        const char* hi = body + len * (k + 1) / nt;
        // a line belongs to the chunk holding its first character
        const char* p = lo;
        if (k > 0 && p[-1] != '\n') {
          while (p < hi && *p != '\n') ++p;
          if (p < hi) ++p;
        }
        while (p < hi) {
          const char* q = p;
          while (q < e && *q != '\n') ++q;
          bool blank = true;
          for (const char* r = p; r < q; ++r)
            if (!std::isspace((unsigned char)*r)) { blank = false; break; }
          if (!blank) starts[k].push_back(p);
          p = q < e ? q + 1 : e;
        }
      });
    for (auto& t : pool) t.join();
  }
  size_t lines = 0;
  for (auto& v : starts) lines += v.size();
  if (lines != n)
    return fail(nullptr, MDPD_ERR_PARSE, "# of lines is not equal to the information specified in file header.");
  std::vector<size_t> first(nt + 1, 0);
  for (unsigned k = 0; k < nt; ++k) first[k + 1] = first[k] + starts[k].size();
  std::vector<int> bad(nt, 0);
  {
    std::vector<std::thread> pool;
    for (unsigned k = 0; k < nt; ++k)
      pool.emplace_back([&, k] {
        for (size_t j = 0; j < starts[k].size(); ++j) {
          const char* p = starts[k][j];
          const char* q = p;
          while (q < e && *q != '\n') ++q;
          if (!xyz_line(p, q, (char*)aos + (first[k] + j) * stride)) bad[k] = 1;
        }
      });
    for (auto& t : pool) t.join();
  }
  for (int x : bad)
    if (x) return fail(nullptr, MDPD_ERR_PARSE, "a particle line does not hold the 17 columns");
  *n_out = n;
  *time_out = time;
  return MDPD_OK;
}

}  // extern "C"
