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
// mdpd_io_wait and mdpd_destroy wait for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"

namespace mdpd {

struct XyzWriter {
  XyzRec* dev = nullptr;   // device staging, upload order
  XyzRec* host = nullptr;  // pinned copy
  uint32_t cap = 0;        // records both buffers hold
  hipEvent_t ev = nullptr;
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

// the writer thread: wait for the copy, format, write
void write_job(XyzWriter* w, std::string path, uint32_t n, uint32_t time, bool append) {
  if (hipEventSynchronize(w->ev) != hipSuccess) {
    w->status = MDPD_ERR_HIP;
    w->err = "trajectory copy failed";
    return;
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
  if (w->ev) (void)hipEventDestroy(w->ev);
  delete w;
}

}  // namespace mdpd

using namespace mdpd;

extern "C" {

int mdpd_write_xyz_async(mdpd_ctx* c, const char* path, uint32_t time, int32_t append) {
  if (!c || !path) return fail(c, MDPD_ERR_ARG, "null argument");
  int rc = join(c);  // the previous write (and its error, if it failed)
  if (rc) return rc;
  if (!c->xyz) {
    c->xyz = new XyzWriter();
    HIPCHK(c, hipEventCreateWithFlags(&c->xyz->ev, hipEventDisableTiming));
  }
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
  HIPCHK(c, hipEventRecord(w->ev, c->st));
  w->status = MDPD_OK;
  w->thr = std::thread(write_job, w, std::string(path), n, time, append != 0);
  return MDPD_OK;
}

int mdpd_io_wait(mdpd_ctx* c) {
  if (!c) return MDPD_ERR_ARG;
  return join(c);
}

}  // extern "C"
