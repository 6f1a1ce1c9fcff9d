// io.hip -- the reference's output and restart formats.
//
//   * the drag / lift history Information/Time_cDcL.dat, Tecplot ASCII
//     (Delay_IO_CD_CL and noDelay_IO_CD_CL, inc/dfib/virtualForceIntergrator.hpp:44-177);
//   * the Plot3D grid file (OutputPlot3D_Xfile, inc/io/Plot3D/xfileWrite.hpp:6-61),
//     solution file (OutputPLOT3D_Qfile, inc/io/Plot3D/qfile.hpp:7-104) and its
//     reader (QfileRead, inc/io/Plot3D/qfileRead.hpp:6-100).
//
// The solution file is assembled on the device: one pass gathers p, the
// cell-centred velocity and eta of every interior cell into the file's
// i-fastest order as float, so the host only copies 5 N floats back and
// writes them.  The grid file, the reader and the text log are host code.
//
// The exact FP64 checkpoint (cfd_checkpoint_save / _load) has no reference
// counterpart: the reference's only restart reads the float Plot3D file
// back, and that block is commented out (RUN_CPU_PC_METHOD.hpp:98-111).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "dev_common.h"

namespace cfd {

namespace {

// One thread per interior cell, o = (k nyc + j) nxc + i (the file order):
// p, 0.5 (u[i-1] + u[i]), 0.5 (v[j-1] + v[j]), 0.5 (w[k-1] + w[k]) and eta,
// each average taken in FP64 and rounded to float once (qfile.hpp:58-98).
__global__ void __launch_bounds__(256) plot3d_pack_kernel(DevGrid g, const double *__restrict__ p,
                                                          const double *__restrict__ u, const double *__restrict__ v,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ eta, float *__restrict__ out) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= g.N) return;
  const int i = (int)(o % g.nxc);
  const int64_t t = o / g.nxc;
  const int j = (int)(t % g.nyc), k = (int)(t / g.nyc);
  const int64_t m = icel(g, i + GC, j + GC, k + GC);
  const int64_t nynz = (int64_t)g.ny * g.nz;
  out[o] = (float)p[m];
  out[g.N + o] = (float)(0.5 * (u[m - nynz] + u[m]));
  out[2 * g.N + o] = (float)(0.5 * (v[m - g.nz] + v[m]));
  out[3 * g.N + o] = (float)(0.5 * (w[m - 1] + w[m]));
  out[4 * g.N + o] = (float)eta[m];
}

struct DevFree {
  void operator()(float *p) const {
    if (p) (void)hipFree(p);
  }
};

// Tecplot header of Time_cDcL.dat (virtualForceIntergrator.hpp:68-88, 144-161)
void cdcl_header(std::ofstream &f, const std::string &zone) {
  f << "TITLE     = \"\"\n"
    << "VARIABLES = \"simulation time\",\"C<sub>D</sub>\",\"C<sub>L</sub>\"\n"
    << "ZONE T=\"" << zone << "\"";
}

bool grid_ok(const cfd_grid_desc *g) {
  return g && g->nx_cal >= 1 && g->ny_cal >= 1 && g->nz_cal >= 1 && g->X && g->Y && g->Z;
}

// ---- checkpoint file -------------------------------------------------------
// Layout (little-endian): CkptHeader, then the arrays of ckpt_arrays() in
// order, then a 64-bit checksum of the arrays' bytes (ckpt_mix over 8-byte
// words, FNV-1a constants).
constexpr char CKPT_MAGIC[8] = {'C', 'F', 'D', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t CKPT_VERSION = 2;

struct CkptHeader {
  char magic[8];
  uint32_t version, abi;
  int32_t nx_cal, ny_cal, nz_cal, loop;
  int32_t solves, solves32, p_from_x, narrays;
  double time;
  uint64_t grid_hash;  // of the extents and the host metric arrays
  uint64_t G, N;
  // bicgstabRe2's counters (bicgstabRe2.hpp:80-84): pResets_, numRestarts_, totalNumIter_
  int32_t nan_resets, restarts;
  int64_t total_iters;
};

constexpr uint64_t FNV_OFFSET = 1469598103934665603ull, FNV_PRIME = 1099511628211ull;

inline uint64_t ckpt_mix(uint64_t h, const void *data, size_t bytes) {
  const unsigned char *b = static_cast<const unsigned char *>(data);
  size_t q = 0;
  for (; q + 8 <= bytes; q += 8) {
    uint64_t w;
    std::memcpy(&w, b + q, 8);
    h = (h ^ w) * FNV_PRIME;
  }
  for (; q < bytes; ++q) h = (h ^ b[q]) * FNV_PRIME;
  return h;
}

uint64_t grid_hash(const Ctx &c) {
  uint64_t h = FNV_OFFSET;
  const int32_t ext[3] = {c.g.nxc, c.g.nyc, c.g.nzc};
  h = ckpt_mix(h, ext, sizeof(ext));
  for (const std::vector<double> *a : {&c.hX, &c.hY, &c.hZ, &c.hDx, &c.hDy, &c.hDz, &c.hDxs, &c.hDys, &c.hDzs})
    h = ckpt_mix(h, a->data(), a->size() * sizeof(double));
  return h;
}

struct CkptArray {
  double *dev;
  size_t len;
};

// what a timestep carries into the next
std::vector<CkptArray> ckpt_arrays(Ctx &c) {
  const size_t G = (size_t)c.g.G, N = (size_t)c.g.N;
  std::vector<CkptArray> a;
  for (int fid = CFD_T0_U; fid <= CFD_ETA; ++fid) a.push_back({c.f[fid], G});  // T0, T1, T3, p, eta
  a.push_back({c.f[CFD_SRC], 3 * N});
  a.push_back({c.f[CFD_X], N});
  a.push_back({c.px1, N});
  a.push_back({c.px2, N});
  return a;
}

constexpr size_t CKPT_CHUNK = (size_t)1 << 21;  // doubles per transfer (16 MiB)

}  // namespace

}  // namespace cfd

using namespace cfd;

// the caller-visible state of simuClass that Delay_IO_CD_CL uses: delayIO and
// the cd / cl ring buffers (simuClass.hpp:11-17, 32), plus the step times for
// the buffered mode
struct cfd_cdcl_log {
  std::string path, zone;
  int delay_io = 100, mode = CFD_CDCL_REFERENCE;
  std::vector<double> t, cd, cl;
};

extern "C" {

int cfd_cdcl_open(const char *path, int delay_io, const char *zone, int mode, cfd_cdcl_log **out) {
  if (!out) return fail(CFD_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!path) return fail(CFD_ERR_ARG, "path is NULL");
  if (delay_io < 1) return fail(CFD_ERR_ARG, "delay_io must be >= 1");
  if (mode != CFD_CDCL_REFERENCE && mode != CFD_CDCL_BUFFERED && mode != CFD_CDCL_EVERY_STEP)
    return fail(CFD_ERR_ARG, "unknown cD/cL log mode");
  auto *log = new (std::nothrow) cfd_cdcl_log;
  if (!log) return fail(CFD_ERR_NOMEM, "cD/cL log allocation failed");
  log->path = path;
  log->zone = zone ? zone : "";
  log->delay_io = delay_io;
  log->mode = mode;
  log->t.assign((size_t)delay_io, 0.0);
  log->cd.assign((size_t)delay_io, 0.0);
  log->cl.assign((size_t)delay_io, 0.0);
  *out = log;
  return CFD_OK;
}

int cfd_cdcl_record(cfd_cdcl_log *log, int loop, double time, double cd, double cl) {
  if (!log) return fail(CFD_ERR_ARG, "log is NULL");
  if (loop < 1) return fail(CFD_ERR_ARG, "loop must be >= 1");
  const char *tab = " ";
  if (log->mode == CFD_CDCL_EVERY_STEP) {  // noDelay_IO_CD_CL (:44-99)
    std::ofstream f(log->path, std::ios::out | std::ios::app);
    if (!f) return fail(CFD_ERR_RUNTIME, "cannot open " + log->path);
    if (loop == 1) cdcl_header(f, log->zone);
    f << "\n" << time << tab << cd << tab << cl;
    return f ? CFD_OK : fail(CFD_ERR_RUNTIME, "write to " + log->path + " failed");
  }
  // Delay_IO_CD_CL (:105-177): ring buffer slot loop % delayIO
  const int D = log->delay_io;
  const int slot = loop % D;
  log->t[(size_t)slot] = time;
  log->cd[(size_t)slot] = cd;
  log->cl[(size_t)slot] = cl;
  if (loop == 1) {
    std::ofstream f(log->path, std::ios::out | std::ios::app);
    if (!f) return fail(CFD_ERR_RUNTIME, "cannot open " + log->path);
    cdcl_header(f, log->zone);
  }
  if (slot != 0) return CFD_OK;
  std::ofstream f(log->path, std::ios::out | std::ios::app);
  if (!f) return fail(CFD_ERR_RUNTIME, "cannot open " + log->path);
  for (int q = 1; q <= D; ++q) {
    // as written, every one of the delayIO lines repeats the current step;
    // the buffered mode writes the steps loop - delayIO + 1 .. loop instead
    const size_t s = log->mode == CFD_CDCL_REFERENCE ? (size_t)slot : (size_t)(q % D);
    const double tv = log->mode == CFD_CDCL_REFERENCE ? time : log->t[s];
    f << "\n" << tv << tab << log->cd[s] << tab << log->cl[s];
  }
  return f ? CFD_OK : fail(CFD_ERR_RUNTIME, "write to " + log->path + " failed");
}

int cfd_cdcl_close(cfd_cdcl_log *log) {
  delete log;
  return CFD_OK;
}

int cfd_plot3d_write_x(const cfd_grid_desc *g, const char *path) {
  if (!grid_ok(g)) return fail(CFD_ERR_ARG, "invalid grid");
  if (!path) return fail(CFD_ERR_ARG, "path is NULL");
  const int nxc = g->nx_cal, nyc = g->ny_cal, nzc = g->nz_cal;
  const size_t N = (size_t)nxc * nyc * nzc;
  // grid::s_init cell centres (structureGrid.hpp:292-305), stored as float
  std::vector<float> xc((size_t)nxc), yc((size_t)nyc), zc((size_t)nzc);
  for (int i = 0; i < nxc; ++i) xc[(size_t)i] = (float)((g->X[i + GC] + g->X[i + GC + 1]) / 2.0);
  for (int j = 0; j < nyc; ++j) yc[(size_t)j] = (float)((g->Y[j + GC] + g->Y[j + GC + 1]) / 2.0);
  for (int k = 0; k < nzc; ++k) zc[(size_t)k] = (float)((g->Z[k + GC] + g->Z[k + GC + 1]) / 2.0);
  std::vector<float> buf(3 * N);
  size_t o = 0;
  for (int k = 0; k < nzc; ++k)
    for (int j = 0; j < nyc; ++j)
      for (int i = 0; i < nxc; ++i, ++o) {
        buf[o] = xc[(size_t)i];
        buf[N + o] = yc[(size_t)j];
        buf[2 * N + o] = zc[(size_t)k];
      }
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(CFD_ERR_RUNTIME, std::string("cannot open ") + path);
  const int hdr[4] = {1, nxc, nyc, nzc};
  const bool ok = std::fwrite(hdr, sizeof(int), 4, f) == 4 && std::fwrite(buf.data(), sizeof(float), 3 * N, f) == 3 * N;
  const bool closed = std::fclose(f) == 0;
  return ok && closed ? CFD_OK : fail(CFD_ERR_RUNTIME, std::string("write to ") + path + " failed");
}

int cfd_plot3d_write_q(cfd_ctx *ctx, const char *path, int vel_field, double time) {
  if (!ctx) return fail(CFD_ERR_ARG, "context is NULL");
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (hipSetDevice(c.device) != hipSuccess) return fail(CFD_ERR_HIP, "hipSetDevice failed");
  if (!path) return fail(CFD_ERR_ARG, "path is NULL");
  if (vel_field != CFD_T0_U && vel_field != CFD_T1_U && vel_field != CFD_T3_U)
    return fail(CFD_ERR_ARG, "vel_field must be CFD_T0_U, CFD_T1_U or CFD_T3_U");
  // t1.p: refreshed from X_result when the last timestep left it there
  if (int rc = sync_pressure(c)) return rc;
  const DevGrid &g = c.g;
  const size_t N = (size_t)g.N;
  float *dbuf = nullptr;
  if (hipMalloc((void **)&dbuf, 5 * N * sizeof(float)) != hipSuccess)
    return fail(CFD_ERR_NOMEM, "hipMalloc of the Plot3D staging buffer failed");
  std::unique_ptr<float, DevFree> hold(dbuf);
  plot3d_pack_kernel<<<blocks_for(g.N, 256, 1 << 30), 256, 0, c.stream>>>(
      g, c.f[CFD_P], c.f[vel_field], c.f[vel_field + 1], c.f[vel_field + 2], c.f[CFD_ETA], dbuf);
  CFD_HIP(hipGetLastError());
  std::vector<float> buf(5 * N);
  CFD_HIP(hipMemcpyAsync(buf.data(), dbuf, 5 * N * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  // header: Nblock, nxCal, nyCal, nzCal, mach 0, alpha 0, Re, time (qfile.hpp:14-55)
  const int ih[4] = {1, g.nxc, g.nyc, g.nzc};
  const float fh[4] = {0.0f, 0.0f, (float)c.prm.Re, (float)time};
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(CFD_ERR_RUNTIME, std::string("cannot open ") + path);
  const bool ok = std::fwrite(ih, sizeof(int), 4, f) == 4 && std::fwrite(fh, sizeof(float), 4, f) == 4 &&
                  std::fwrite(buf.data(), sizeof(float), 5 * N, f) == 5 * N;
  const bool closed = std::fclose(f) == 0;
  return ok && closed ? CFD_OK : fail(CFD_ERR_RUNTIME, std::string("write to ") + path + " failed");
}

int cfd_plot3d_read_q(const char *path, const cfd_grid_desc *g, int *nblock, float header[4], double *p, double *u,
                      double *v, double *w, double *eta) {
  if (!grid_ok(g)) return fail(CFD_ERR_ARG, "invalid grid");
  if (!path || !nblock || !header || !p || !u || !v || !w || !eta) return fail(CFD_ERR_ARG, "NULL argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(CFD_ERR_RUNTIME, std::string("cannot read ") + path);
  int ih[4];
  float fh[4];
  bool ok = std::fread(ih, sizeof(int), 4, f) == 4 && std::fread(fh, sizeof(float), 4, f) == 4;
  if (ok && (ih[1] != g->nx_cal || ih[2] != g->ny_cal || ih[3] != g->nz_cal)) {
    std::fclose(f);
    return fail(CFD_ERR_ARG, std::string(path) + " holds a " + std::to_string(ih[1]) + "x" + std::to_string(ih[2]) +
                                 "x" + std::to_string(ih[3]) + " grid, not the given one");
  }
  const int nxc = g->nx_cal, nyc = g->ny_cal, nzc = g->nz_cal;
  const size_t N = (size_t)nxc * nyc * nzc;
  std::vector<float> raw(5 * N);
  ok = ok && std::fread(raw.data(), sizeof(float), 5 * N, f) == 5 * N;
  std::fclose(f);
  if (!ok) return fail(CFD_ERR_RUNTIME, std::string(path) + " is truncated");
  *nblock = ih[0];
  for (int q = 0; q < 4; ++q) header[q] = fh[q];
  // data[q][(i nyc + j) nzc + k] from the file's i-fastest order
  std::vector<float> data(5 * N);
  for (int q = 0; q < 5; ++q) {
    size_t o = 0;
    for (int k = 0; k < nzc; ++k)
      for (int j = 0; j < nyc; ++j)
        for (int i = 0; i < nxc; ++i, ++o) data[q * N + ((size_t)i * nyc + j) * nzc + k] = raw[q * N + o];
  }
  const int ny = nyc + 2 * GC, nz = nzc + 2 * GC;
  const size_t G = (size_t)(nxc + 2 * GC) * ny * nz;
  double *out[5] = {p, u, v, w, eta};
  for (double *a : out) std::fill(a, a + G, 0.0);
  const float *d0 = data.data(), *d1 = d0 + N, *d2 = d1 + N, *d3 = d2 + N, *d4 = d3 + N;
  for (int i = 0; i < nxc; ++i)
    for (int j = 0; j < nyc; ++j)
      for (int k = 0; k < nzc; ++k) {
        const size_t m = ((size_t)(i + GC) * ny + (j + GC)) * nz + (k + GC);
        const size_t cc = ((size_t)i * nyc + j) * nzc + k;
        p[m] = d0[cc];
        eta[m] = d4[cc];
        // the face value is the mean of the two cell-centred floats, summed
        // in float and halved in double (qfileRead.hpp:73-94); the last
        // cell of an axis keeps its own value
        u[m] = i == nxc - 1 ? (double)d1[cc] : (d1[cc] + d1[cc + (size_t)nyc * nzc]) / 2.0;
        v[m] = j == nyc - 1 ? (double)d2[cc] : (d2[cc] + d2[cc + nzc]) / 2.0;
        w[m] = k == nzc - 1 ? (double)d3[cc] : (d3[cc] + d3[cc + 1]) / 2.0;
      }
  return CFD_OK;
}

int cfd_checkpoint_save(cfd_ctx *ctx, const char *path) {
  if (!ctx) return fail(CFD_ERR_ARG, "context is NULL");
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (hipSetDevice(c.device) != hipSuccess) return fail(CFD_ERR_HIP, "hipSetDevice failed");
  if (!path) return fail(CFD_ERR_ARG, "path is NULL");
  const std::vector<CkptArray> arrays = ckpt_arrays(c);
  CkptHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, CKPT_MAGIC, 8);
  h.version = CKPT_VERSION;
  h.abi = CFD_ABI_VERSION;
  h.nx_cal = c.g.nxc;
  h.ny_cal = c.g.nyc;
  h.nz_cal = c.g.nzc;
  h.loop = c.loop;
  h.solves = c.solves;
  h.solves32 = c.solves32;
  h.p_from_x = c.p_from_x ? 1 : 0;
  h.narrays = (int32_t)arrays.size();
  h.time = c.time;
  h.grid_hash = grid_hash(c);
  h.G = (uint64_t)c.g.G;
  h.N = (uint64_t)c.g.N;
  CFD_HIP(hipStreamSynchronize(c.stream));  // the state of every queued step
  KrylovScalars ks;
  CFD_HIP(hipMemcpy(&ks, c.ks, sizeof(ks), hipMemcpyDeviceToHost));
  h.nan_resets = c.nan_resets;
  h.restarts = ks.restarts;
  h.total_iters = c.total_iters;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(CFD_ERR_RUNTIME, std::string("cannot open ") + path);
  std::unique_ptr<FILE, int (*)(FILE *)> hold(f, std::fclose);
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  uint64_t sum = FNV_OFFSET;
  std::vector<double> buf(CKPT_CHUNK);
  for (const CkptArray &a : arrays) {
    for (size_t off = 0; ok && off < a.len; off += CKPT_CHUNK) {
      const size_t n = std::min(CKPT_CHUNK, a.len - off);
      CFD_HIP(hipMemcpyAsync(buf.data(), a.dev + off, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      CFD_HIP(hipStreamSynchronize(c.stream));
      sum = ckpt_mix(sum, buf.data(), n * sizeof(double));
      ok = std::fwrite(buf.data(), sizeof(double), n, f) == n;
    }
  }
  ok = ok && std::fwrite(&sum, sizeof(sum), 1, f) == 1;
  const bool closed = std::fclose(hold.release()) == 0;
  return ok && closed ? CFD_OK : fail(CFD_ERR_RUNTIME, std::string("write to ") + path + " failed");
}

int cfd_checkpoint_load(cfd_ctx *ctx, const char *path) {
  if (!ctx) return fail(CFD_ERR_ARG, "context is NULL");
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (hipSetDevice(c.device) != hipSuccess) return fail(CFD_ERR_HIP, "hipSetDevice failed");
  if (!path) return fail(CFD_ERR_ARG, "path is NULL");
  const std::vector<CkptArray> arrays = ckpt_arrays(c);
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(CFD_ERR_RUNTIME, std::string("cannot open ") + path);
  std::unique_ptr<FILE, int (*)(FILE *)> hold(f, std::fclose);
  CkptHeader h;
  if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, CKPT_MAGIC, 8) != 0)
    return fail(CFD_ERR_RUNTIME, std::string(path) + " is not a checkpoint");
  if (h.version != CKPT_VERSION || h.narrays != (int32_t)arrays.size())
    return fail(CFD_ERR_RUNTIME, std::string(path) + ": unsupported checkpoint version");
  if (h.nx_cal != c.g.nxc || h.ny_cal != c.g.nyc || h.nz_cal != c.g.nzc || h.G != (uint64_t)c.g.G ||
      h.N != (uint64_t)c.g.N || h.grid_hash != grid_hash(c))
    return fail(CFD_ERR_ARG, std::string(path) + " was written for another grid");
  if (h.loop < 1 || h.solves < 0 || h.solves32 < 0 || !std::isfinite(h.time) || h.nan_resets < 0 ||
      h.restarts < 0 || h.total_iters < 0)
    return fail(CFD_ERR_RUNTIME, std::string(path) + ": invalid step counters");
  // first pass: the whole file is read and its checksum verified before any
  // of the context's state is touched
  const long data_at = std::ftell(f);
  std::vector<double> buf(CKPT_CHUNK);
  uint64_t sum = FNV_OFFSET;
  for (const CkptArray &a : arrays)
    for (size_t off = 0; off < a.len; off += CKPT_CHUNK) {
      const size_t n = std::min(CKPT_CHUNK, a.len - off);
      if (std::fread(buf.data(), sizeof(double), n, f) != n)
        return fail(CFD_ERR_RUNTIME, std::string(path) + " is truncated");
      sum = ckpt_mix(sum, buf.data(), n * sizeof(double));
    }
  uint64_t stored = 0;
  if (std::fread(&stored, sizeof(stored), 1, f) != 1 || stored != sum)
    return fail(CFD_ERR_RUNTIME, std::string(path) + ": checksum mismatch");
  if (std::fgetc(f) != EOF) return fail(CFD_ERR_RUNTIME, std::string(path) + ": trailing bytes");
  // second pass: the arrays to the device
  if (std::fseek(f, data_at, SEEK_SET) != 0) return fail(CFD_ERR_RUNTIME, std::string("cannot seek in ") + path);
  CFD_HIP(hipStreamSynchronize(c.stream));  // nothing queued still reads or writes these
  for (const CkptArray &a : arrays)
    for (size_t off = 0; off < a.len; off += CKPT_CHUNK) {
      const size_t n = std::min(CKPT_CHUNK, a.len - off);
      if (std::fread(buf.data(), sizeof(double), n, f) != n)
        return fail(CFD_ERR_RUNTIME, std::string(path) + " changed while it was read");
      CFD_HIP(hipMemcpyAsync(a.dev + off, buf.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream));
      CFD_HIP(hipStreamSynchronize(c.stream));  // buf is refilled next
    }
  c.loop = h.loop;
  c.time = h.time;
  c.solves = h.solves;
  c.solves32 = h.solves32;
  c.p_from_x = h.p_from_x != 0;
  c.nan_resets = h.nan_resets;
  c.total_iters = h.total_iters;
  KrylovScalars ks;
  CFD_HIP(hipMemcpy(&ks, c.ks, sizeof(ks), hipMemcpyDeviceToHost));
  ks.restarts = h.restarts;
  CFD_HIP(hipMemcpy(c.ks, &ks, sizeof(ks), hipMemcpyHostToDevice));
  adjoint_invalidate(c);
  return CFD_OK;
}

}  // extern "C"
