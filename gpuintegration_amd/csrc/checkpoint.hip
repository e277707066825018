// checkpoint.hip -- region-pool checkpoints and the CSV region dump.
//
// The reference keeps one in-memory copy of a region list
// (Sub_regions::take_snapshot / load_snapshot, Sub_regions.cuh:220-239) and
// writes region lists as CSV (Snapshot<NDIM>::Save, quad.h:78-152).  Here a
// run suspended at an iteration boundary (pagani_ws_set_iteration_limit) can
// be written to a file and continued later, in this process or another,
// with exactly the result the uninterrupted run gives: the file holds the
// whole loop state (the region list, the parents' estimates, the running
// sums, the classifier's history and the per-iteration trace).
//
// File layout, little-endian, version 2 (all integers unsigned unless noted):
//   "PAGANIC1"  u32 version  u32 ndim
//   u64 n  u64 it  i32 has_parents  [i32 split, version 2]  i32 user_regions  i32 relerr
//   i32 mem_model
//   u64 mem_total  f64 epsrel  f64 epsabs  f64 lows[16]  f64 highs[16]
//   result: f64 estimate  f64 errorest  u64 neval  u64 nregions
//           u64 nFinishedRegions  i32 status  i32 lastPhase  f64 chi_sq  u64 iters
//   f64 hist[3]  u64 iters_collected
//   i32 family  i32 genz_ndim  f64 a[16]  f64 u[16]  i32 precision
//   u64 ntrace, then ntrace x (u64 n, u64 m, f64 estimate, f64 errorest)
//   f64 L[ndim*rows]  f64 H[ndim*rows]  [u8 pdim[rows] when split]
//   f64 parent_est[has_parents ? n/2 : 0]
//   where rows = n for a flat list and n/2 for a list in split form (the
//   parents' bounds and the axis each is halved along; region t is child
//   t / rows of parent t mod rows).  Version 1 files (no split field, the
//   per-axis arrays 8 wide) hold flat lists and still load.
//   u64 checksum of everything before it (FNV-1a over 64-bit words, the
//       byte tail zero-padded)
// Arrays are structure of arrays, [j*n + t], unit-cube coordinates.

#include <hip/hip_runtime.h>

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ws_internal.hpp"

using namespace pagani::internal;

namespace {

constexpr char kMagic[8] = {'P', 'A', 'G', 'A', 'N', 'I', 'C', '1'};
constexpr std::uint32_t kVersion = 2;
constexpr std::size_t kStageBytes = std::size_t(64) << 20;

// FNV-1a over 64-bit words; a partial word at the end of a write is held
// back until the next one (or zero-padded by finish()).
struct Checksum {
  std::uint64_t h = 0xcbf29ce484222325ull;
  unsigned char tail[8];
  int ntail = 0;
  void word(std::uint64_t w) { h = (h ^ w) * 0x100000001b3ull; }
  void add(const void* p, std::size_t len)
  {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    while (len && ntail) {
      tail[ntail++] = *b++;
      --len;
      if (ntail == 8) {
        std::uint64_t w;
        std::memcpy(&w, tail, 8);
        word(w);
        ntail = 0;
      }
    }
    for (; len >= 8; len -= 8, b += 8) {
      std::uint64_t w;
      std::memcpy(&w, b, 8);
      word(w);
    }
    for (; len; --len) tail[ntail++] = *b++;
  }
  std::uint64_t finish() const
  {
    Checksum c = *this;
    if (c.ntail) {
      unsigned char t[8] = {};
      std::memcpy(t, c.tail, c.ntail);
      std::uint64_t w;
      std::memcpy(&w, t, 8);
      c.word(w);
    }
    return c.h;
  }
};

struct Writer {
  std::FILE* f;
  Checksum ck;
  bool ok = true;
  void raw(const void* p, std::size_t len)
  {
    if (!ok || len == 0) return;
    ok = std::fwrite(p, 1, len, f) == len;
    ck.add(p, len);
  }
  template <class T>
  void put(T v)
  {
    raw(&v, sizeof v);
  }
};

struct Reader {
  std::FILE* f;
  Checksum ck;
  bool ok = true;
  void raw(void* p, std::size_t len)
  {
    if (!ok || len == 0) return;
    ok = std::fread(p, 1, len, f) == len;
    if (ok) ck.add(p, len);
  }
  template <class T>
  T get()
  {
    T v{};
    raw(&v, sizeof v);
    return v;
  }
};

struct Stage {
  void* p = nullptr;
  ~Stage()
  {
    if (p) (void)hipHostFree(p);
  }
};

template <class T>
int write_device_array(pagani_ws* ws, Writer& w, const T* d, std::size_t count, Stage& st)
{
  const std::size_t per = kStageBytes / sizeof(T);
  for (std::size_t i = 0; i < count && w.ok; i += per) {
    const std::size_t c = std::min(per, count - i);
    HIP_TRY(ws, hipMemcpyAsync(st.p, d + i, c * sizeof(T), hipMemcpyDeviceToHost, ws->stream));
    HIP_TRY(ws, hipStreamSynchronize(ws->stream));
    w.raw(st.p, c * sizeof(T));
  }
  return w.ok ? PAGANI_OK : fail(ws, PAGANI_EBADARG, "checkpoint write failed");
}

// The widths of a list held as levels (pagani_eval_args), written as the
// doubles a list with double widths holds: width = h0[j] * 2^-level, row-major
// per axis like the levels.
int write_level_widths(pagani_ws* ws, Writer& w, const std::uint16_t* d, std::size_t rows,
                       std::size_t ndim, const pagani::phases::Levels& lv, Stage& st)
{
  const std::size_t per = kStageBytes / sizeof(double);
  std::vector<double> wd(per);
  for (std::size_t j = 0; j < ndim && w.ok; ++j)
    for (std::size_t i = 0; i < rows && w.ok; i += per) {
      const std::size_t c = std::min(per, rows - i);
      HIP_TRY(ws, hipMemcpyAsync(st.p, d + j * rows + i, c * sizeof(std::uint16_t),
                                 hipMemcpyDeviceToHost, ws->stream));
      HIP_TRY(ws, hipStreamSynchronize(ws->stream));
      const std::uint16_t* lev = static_cast<const std::uint16_t*>(st.p);
      for (std::size_t k = 0; k < c; ++k) wd[k] = std::ldexp(lv.h0[j], -static_cast<int>(lev[k]));
      w.raw(wd.data(), c * sizeof(double));
    }
  return w.ok ? PAGANI_OK : fail(ws, PAGANI_EBADARG, "checkpoint write failed");
}

template <class T>
int read_device_array(pagani_ws* ws, Reader& r, T* d, std::size_t count, Stage& st)
{
  const std::size_t per = kStageBytes / sizeof(T);
  for (std::size_t i = 0; i < count && r.ok; i += per) {
    const std::size_t c = std::min(per, count - i);
    r.raw(st.p, c * sizeof(T));
    if (!r.ok) break;
    HIP_TRY(ws, hipMemcpyAsync(d + i, st.p, c * sizeof(T), hipMemcpyHostToDevice, ws->stream));
    HIP_TRY(ws, hipStreamSynchronize(ws->stream));
  }
  return r.ok ? PAGANI_OK : fail(ws, PAGANI_EBADARG, "checkpoint truncated");
}

struct FileCloser {
  std::FILE* f;
  ~FileCloser()
  {
    if (f) std::fclose(f);
  }
};

}  // namespace

extern "C" {

int pagani_ws_save_state(pagani_ws* ws, const char* path)
{
  if (!ws || !path) return PAGANI_EBADARG;
  if (!ws->susp.held) return fail(ws, PAGANI_EBADARG, "no suspended run to save");
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  const LoopState& s = ws->susp;
  const std::size_t d = static_cast<std::size_t>(ws->ndim);
  const std::size_t np = s.has_parents ? s.n / 2 : 0;
  const std::size_t rows = s.split ? s.n / 2 : s.n;
  const std::string tmp = std::string(path) + ".tmp";
  FileCloser fc{std::fopen(tmp.c_str(), "wb")};
  if (!fc.f) return fail(ws, PAGANI_EBADARG, std::string("cannot open ") + tmp);
  Writer w{fc.f};
  w.raw(kMagic, sizeof kMagic);
  w.put<std::uint32_t>(kVersion);
  w.put<std::uint32_t>(static_cast<std::uint32_t>(d));
  w.put<std::uint64_t>(s.n);
  w.put<std::uint64_t>(s.it);
  w.put<std::int32_t>(s.has_parents);
  w.put<std::int32_t>(s.split);
  w.put<std::int32_t>(s.user_regions);
  w.put<std::int32_t>(s.relerr);
  w.put<std::int32_t>(s.mem_model);
  w.put<std::uint64_t>(s.mem_total);
  w.put<double>(s.epsrel);
  w.put<double>(s.epsabs);
  for (int j = 0; j < PAGANI_MAX_DIM; ++j) w.put<double>(s.lows[j]);
  for (int j = 0; j < PAGANI_MAX_DIM; ++j) w.put<double>(s.highs[j]);
  w.put<double>(s.cum.estimate);
  w.put<double>(s.cum.errorest);
  w.put<std::uint64_t>(s.cum.neval);
  w.put<std::uint64_t>(s.cum.nregions);
  w.put<std::uint64_t>(s.cum.nFinishedRegions);
  w.put<std::int32_t>(s.cum.status);
  w.put<std::int32_t>(s.cum.lastPhase);
  w.put<double>(s.cum.chi_sq);
  w.put<std::uint64_t>(s.cum.iters);
  for (double h : s.hist) w.put<double>(h);
  w.put<std::uint64_t>(s.iters_collected);
  w.put<std::int32_t>(s.genz.family);
  w.put<std::int32_t>(s.genz.ndim);
  for (int j = 0; j < PAGANI_MAX_DIM; ++j) w.put<double>(s.genz.a[j]);
  for (int j = 0; j < PAGANI_MAX_DIM; ++j) w.put<double>(s.genz.u[j]);
  w.put<std::int32_t>(s.precision);
  w.put<std::uint64_t>(ws->trace.size());
  for (const TraceRec& t : ws->trace) {
    w.put<std::uint64_t>(t.n);
    w.put<std::uint64_t>(t.m);
    w.put<double>(t.estimate);
    w.put<double>(t.errorest);
  }
  Stage st;
  if (hipHostMalloc(&st.p, kStageBytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return fail(ws, PAGANI_ENOMEM, "pinned staging buffer");
  }
  int rc;
  // widths held as levels go to the file as doubles: a loaded run uses
  // double widths
  if ((rc = write_device_array(ws, w, ws->L[ws->cur].p, d * rows, st)) ||
      (rc = s.lv.on ? write_level_widths(ws, w, reinterpret_cast<const std::uint16_t*>(
                                                    ws->H[ws->cur].p),
                                         rows, d, s.lv, st)
                    : write_device_array(ws, w, ws->H[ws->cur].p, d * rows, st)) ||
      (s.split && (rc = write_device_array(ws, w, ws->pdim[ws->cur].p, rows, st))) ||
      (rc = write_device_array(ws, w, ws->pest.p, np, st)))
    return rc;
  const std::uint64_t sum = w.ck.finish();
  if (w.ok) w.ok = std::fwrite(&sum, 1, sizeof sum, fc.f) == sizeof sum;
  if (!w.ok) return fail(ws, PAGANI_EBADARG, "checkpoint write failed");
  if (std::fclose(fc.f) != 0) {
    fc.f = nullptr;
    return fail(ws, PAGANI_EBADARG, "checkpoint close failed");
  }
  fc.f = nullptr;
  if (std::rename(tmp.c_str(), path) != 0)
    return fail(ws, PAGANI_EBADARG, std::string("cannot rename to ") + path);
  return PAGANI_OK;
}

int pagani_ws_load_state(pagani_ws* ws, const char* path)
{
  if (!ws || !path) return PAGANI_EBADARG;
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  FileCloser fc{std::fopen(path, "rb")};
  if (!fc.f) return fail(ws, PAGANI_EBADARG, std::string("cannot open ") + path);
  std::fseek(fc.f, 0, SEEK_END);
  const long fsize = std::ftell(fc.f);
  std::fseek(fc.f, 0, SEEK_SET);
  Reader r{fc.f};
  char magic[8];
  r.raw(magic, sizeof magic);
  if (!r.ok || std::memcmp(magic, kMagic, sizeof magic) != 0)
    return fail(ws, PAGANI_EBADARG, "not a PAGANI checkpoint");
  const std::uint32_t version = r.get<std::uint32_t>();
  if (version != 1 && version != kVersion)
    return fail(ws, PAGANI_EBADARG, "unsupported checkpoint version");
  const std::uint32_t d = r.get<std::uint32_t>();
  if (!r.ok || static_cast<int>(d) != ws->ndim)
    return fail(ws, PAGANI_EBADARG, "checkpoint ndim differs from the workspace's");
  LoopState s;
  s.n = r.get<std::uint64_t>();
  s.it = r.get<std::uint64_t>();
  s.has_parents = r.get<std::int32_t>();
  s.split = version >= 2 ? r.get<std::int32_t>() : 0;
  s.user_regions = r.get<std::int32_t>();
  s.relerr = r.get<std::int32_t>();
  s.mem_model = r.get<std::int32_t>();
  s.mem_total = r.get<std::uint64_t>();
  s.epsrel = r.get<double>();
  s.epsabs = r.get<double>();
  // per-axis arrays: 8 wide in version 1, PAGANI_MAX_DIM (16) from version 2
  const int axes = version >= 2 ? PAGANI_MAX_DIM : 8;
  for (int j = 0; j < axes; ++j) s.lows[j] = r.get<double>();
  for (int j = 0; j < axes; ++j) s.highs[j] = r.get<double>();
  s.cum.estimate = r.get<double>();
  s.cum.errorest = r.get<double>();
  s.cum.neval = r.get<std::uint64_t>();
  s.cum.nregions = r.get<std::uint64_t>();
  s.cum.nFinishedRegions = r.get<std::uint64_t>();
  s.cum.status = r.get<std::int32_t>();
  s.cum.lastPhase = r.get<std::int32_t>();
  s.cum.chi_sq = r.get<double>();
  s.cum.iters = r.get<std::uint64_t>();
  for (double& h : s.hist) h = r.get<double>();
  s.iters_collected = r.get<std::uint64_t>();
  s.genz.family = r.get<std::int32_t>();
  s.genz.ndim = r.get<std::int32_t>();
  for (int j = 0; j < axes; ++j) s.genz.a[j] = r.get<double>();
  for (int j = 0; j < axes; ++j) s.genz.u[j] = r.get<double>();
  s.precision = r.get<std::int32_t>();
  const std::uint64_t ntrace = r.get<std::uint64_t>();
  if (!r.ok) return fail(ws, PAGANI_EBADARG, "checkpoint truncated");
  // sizes must match the file before anything is allocated
  const std::uint64_t np = s.has_parents ? s.n / 2 : 0;
  const std::uint64_t rows = s.split ? s.n / 2 : s.n;
  const long head = std::ftell(fc.f);
  const double want = static_cast<double>(head) + 32.0 * static_cast<double>(ntrace) +
                      8.0 * (2.0 * d * static_cast<double>(rows) + static_cast<double>(np)) +
                      (s.split ? static_cast<double>(rows) : 0.0) + 8.0;
  if (fsize < 0 || static_cast<double>(fsize) != want || ((s.has_parents || s.split) && (s.n & 1)) ||
      s.n == 0 || (s.has_parents != 0 && s.has_parents != 1) || (s.split != 0 && s.split != 1))
    return fail(ws, PAGANI_EBADARG, "checkpoint size or header inconsistent");
  std::vector<TraceRec> trace(ntrace);
  for (TraceRec& t : trace) {
    t.n = r.get<std::uint64_t>();
    t.m = r.get<std::uint64_t>();
    t.estimate = r.get<double>();
    t.errorest = r.get<double>();
  }
  int rc;
  const int cur = ws->cur;
  // From here on the live buffers are overwritten (and may be reallocated):
  // a run the workspace held is gone, whether or not this load succeeds, so
  // a failed load leaves nothing that a resume could continue.
  ws->susp.held = false;
  ws->live_n = 0;
  ws->live_evaluated = false;
  ws->snap_n = 0;
  ws->trace.clear();
  if ((rc = ensure(ws, ws->L[cur], d * rows, "region bounds")) ||
      (rc = ensure(ws, ws->H[cur], d * rows, "region bounds")) ||
      (s.split && (rc = ensure(ws, ws->pdim[cur], rows, "split axes"))) ||
      (np && (rc = ensure(ws, ws->pest, np, "parent estimates"))))
    return rc;
  Stage st;
  if (hipHostMalloc(&st.p, kStageBytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return fail(ws, PAGANI_ENOMEM, "pinned staging buffer");
  }
  if ((rc = read_device_array(ws, r, ws->L[cur].p, d * rows, st)) ||
      (rc = read_device_array(ws, r, ws->H[cur].p, d * rows, st)) ||
      (s.split && (rc = read_device_array(ws, r, ws->pdim[cur].p, rows, st))) ||
      (rc = read_device_array(ws, r, ws->pest.p, np, st)))
    return rc;
  const std::uint64_t expect = r.ck.finish();
  std::uint64_t sum = 0;
  if (std::fread(&sum, 1, sizeof sum, fc.f) != sizeof sum || sum != expect)
    return fail(ws, PAGANI_EBADARG, "checkpoint checksum mismatch");
  s.held = true;
  ws->susp = s;
  ws->trace = std::move(trace);
  ws->live_n = s.n;
  ws->live_evaluated = false;
  ws->split = s.split != 0;
  ws->lv = pagani::phases::Levels{};  // the file holds double widths
  return PAGANI_OK;
}

// The live region list as CSV rows in the column order of the reference's
// Snapshot<NDIM>::Save (quad.h:107-118): estimate, error, then upper and
// lower bound per axis, the split axis and -1.  Values are written with 17
// significant digits; bounds are in unit-cube coordinates.  A list that has
// not been evaluated (a suspended run) has nan estimates and split axis -1.
int pagani_ws_write_regions_csv(pagani_ws* ws, const char* path)
{
  if (!ws || !path) return PAGANI_EBADARG;
  if (hipSetDevice(ws->device) != hipSuccess) return fail(ws, PAGANI_EHIP, "hipSetDevice");
  const std::size_t n = ws->live_n, d = static_cast<std::size_t>(ws->ndim);
  // a list in split form is read as its parents and written out child by
  // child, each derived as the evaluation derives it
  const std::size_t rows = ws->split ? n / 2 : n;
  std::vector<double> L(rows * d), H(rows * d), est, err;
  std::vector<std::uint8_t> dim, pdim(ws->split ? rows : 0);
  if (n) {
    HIP_TRY(ws, hipMemcpy(L.data(), ws->L[ws->cur].p, rows * d * sizeof(double),
                          hipMemcpyDeviceToHost));
    if (ws->lv.on) {
      std::vector<std::uint16_t> lev(rows * d);
      HIP_TRY(ws, hipMemcpy(lev.data(), ws->H[ws->cur].p, rows * d * sizeof(std::uint16_t),
                            hipMemcpyDeviceToHost));
      for (std::size_t j = 0; j < d; ++j)
        for (std::size_t i = 0; i < rows; ++i)
          H[j * rows + i] = std::ldexp(ws->lv.h0[j], -static_cast<int>(lev[j * rows + i]));
    } else {
      HIP_TRY(ws, hipMemcpy(H.data(), ws->H[ws->cur].p, rows * d * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    if (ws->split)
      HIP_TRY(ws, hipMemcpy(pdim.data(), ws->pdim[ws->cur].p, rows, hipMemcpyDeviceToHost));
    if (ws->live_evaluated) {
      est.resize(n);
      err.resize(n);
      dim.resize(n);
      HIP_TRY(ws, hipMemcpy(est.data(), ws->est.p, n * sizeof(double), hipMemcpyDeviceToHost));
      HIP_TRY(ws, hipMemcpy(err.data(), ws->live_err1 ? ws->err1.p : ws->err0.p,
                            n * sizeof(double), hipMemcpyDeviceToHost));
      HIP_TRY(ws, hipMemcpy(dim.data(), ws->dim.p, n, hipMemcpyDeviceToHost));
    }
  }
  FileCloser fc{std::fopen(path, "w")};
  if (!fc.f) return fail(ws, PAGANI_EBADARG, std::string("cannot open ") + path);
  bool ok = true;
  for (std::size_t t = 0; t < n && ok; ++t) {
    if (ws->live_evaluated)
      ok = std::fprintf(fc.f, "%.17g,%.17g,", est[t], err[t]) > 0;
    else
      ok = std::fprintf(fc.f, "nan,nan,") > 0;
    const bool right = ws->split && t >= rows;
    const std::size_t row = right ? t - rows : t;
    for (std::size_t j = 0; j < d && ok; ++j) {
      double lo = L[j * rows + row], len = H[j * rows + row];
      if (ws->split && j == pdim[row]) {
        const double half = len / 2;
        if (right) lo = lo + half;
        len = half;
      }
      const double up = lo + len;
      ok = std::fprintf(fc.f, "%.17g,%.17g,", up, lo) > 0;
    }
    if (ok)
      ok = std::fprintf(fc.f, "%d,-1\n", ws->live_evaluated ? static_cast<int>(dim[t]) : -1) > 0;
  }
  if (!ok) return fail(ws, PAGANI_EBADARG, "csv write failed");
  return PAGANI_OK;
}

}  // extern "C"
