// C ABI: context lifecycle, resident vector store, reference-shaped drop-ins
// and the device engine entry points.
#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "mjk_internal.h"

namespace mjk {

int set_error(Context *c, int rc, const std::string &msg) {
  if (c) c->err = msg;
  return rc;
}

int hip_check(Context *c, hipError_t e, const char *what) {
  if (e == hipSuccess) return MJK_OK;
  const int rc = (e == hipErrorOutOfMemory) ? MJK_ERR_NOMEM : MJK_ERR_HIP;
  return set_error(c, rc, std::string(what) + ": " + hipGetErrorString(e));
}

int ensure(Context *c, DevBuf &b, size_t bytes) {
  if (bytes <= b.bytes && b.ptr) return MJK_OK;
  if (b.ptr) {
    // queued work may still read the old buffer
    int rc = hip_check(c, hipStreamSynchronize(c->stream), "sync before realloc");
    if (rc) return rc;
    hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  const size_t want = std::max<size_t>(bytes, 256);
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(c, MJK_ERR_NOMEM,
                     "device allocation of " + std::to_string(want) + " bytes failed");
  }
  b.ptr = p;
  b.bytes = want;
  return MJK_OK;
}

namespace {
// Zero fill in 16-byte words plus a 4-byte tail (all buffers are 4-byte multiples).
__global__ void k_fill_zero(uint4 *__restrict__ p16, int64_t n16, uint32_t *__restrict__ p4,
                            int64_t n4) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
    p16[i] = make_uint4(0, 0, 0, 0);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += stride) p4[i] = 0u;
}
}  // namespace

int fill_zero(Context *c, void *ptr, size_t bytes) {
  if (bytes == 0) return MJK_OK;
  if ((reinterpret_cast<uintptr_t>(ptr) & 3) != 0 || (bytes & 3) != 0)
    return set_error(c, MJK_ERR_ARG, "fill_zero: buffer not 4-byte aligned");
  // 4-byte head up to 16-byte alignment, 16-byte body, 4-byte tail
  char *p = static_cast<char *>(ptr);
  const size_t head = std::min(bytes, (16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15);
  const size_t body = (bytes - head) / 16 * 16;
  const size_t tail = bytes - head - body;
  const int64_t n16 = int64_t(body / 16);
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(4096, ceil_div(n16, 256)));
  if (head)
    hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, c->stream, static_cast<uint4 *>(nullptr),
                       int64_t(0), reinterpret_cast<uint32_t *>(p), int64_t(head / 4));
  hipLaunchKernelGGL(k_fill_zero, dim3(unsigned(blocks)), dim3(256), 0, c->stream,
                     reinterpret_cast<uint4 *>(p + head), n16,
                     reinterpret_cast<uint32_t *>(p + head + body), int64_t(tail / 4));
  return hip_check(c, hipGetLastError(), "fill_zero");
}

static hipEvent_t take_event(Context *c, int *idx) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  *idx = int(c->ev_used);
  return c->ev_pool[c->ev_used++];
}

void phase_begin(Context *c) {
  if (!c->profile) return;
  int idx;
  hipEvent_t e = take_event(c, &idx);
  if (!e) return;
  hipEventRecord(e, c->stream);
  c->ev_open = idx;
}

void phase_end(Context *c, int phase) {
  if (!c->profile || c->ev_open < 0) return;
  int idx;
  hipEvent_t e = take_event(c, &idx);
  if (!e) return;
  hipEventRecord(e, c->stream);
  c->pending.push_back({phase, c->ev_open, idx});
  c->ev_open = -1;
  if (c->ev_used > 8192) resolve_phases(c);  // bound the pool
}

int resolve_phases(Context *c) {
  if (c->pending.empty()) {
    c->ev_used = 0;
    return MJK_OK;
  }
  int rc = hip_check(c, hipEventSynchronize(c->ev_pool[c->pending.back().ev_end]), "event sync");
  if (rc) return rc;
  for (const auto &p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_pool[p.ev_begin], c->ev_pool[p.ev_end]) == hipSuccess) {
      c->phase_ms[p.phase] += ms;
      c->phase_calls[p.phase] += 1;
    }
  }
  c->pending.clear();
  c->ev_used = 0;
  c->ev_open = -1;
  return MJK_OK;
}

namespace {

// L[(jvec + j)*ldl + map[r]] = src[r + j*nrs]; grid-stride over vectors (y) and
// reduced-set positions (x), so a launch of any size fits the grid limits.
__global__ void k_scatter_rs(const double *__restrict__ src, int64_t nrs, int64_t jnum,
                             const int32_t *__restrict__ map, double *__restrict__ L, int64_t ldl,
                             int64_t jvec) {
  for (int64_t j = blockIdx.y; j < jnum; j += gridDim.y)
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrs;
         r += int64_t(gridDim.x) * blockDim.x)
      L[(jvec + j) * ldl + map[r]] = src[r + j * nrs];
}

void free_buf(DevBuf &b) {
  if (b.ptr) hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

}  // namespace
}  // namespace mjk

using namespace mjk;

// LK thresholds of the drop-ins, per density: the basic threshold of
// Cho_LK_ScreeningThreshold (cho_lk_screeningthreshold.F90:34-39) and
//   tau  = LKThr / max(1,nOcc_d) * dmpk  (/ MaxRed with Estimate)  (cho_lk_scf.F90:155-165)
//   thrv = sqrt(LKThr / (max(1,nOcc_d) * NumVT)) * dmpk            (cho_lk_scf.F90:170-171)
// (cho_lk_casscf.F90:181-200 with the inactive / active orbital counts).
static void lk_thresholds(const mjk_context *c, int nden, const int32_t *nocc, double dmpk,
                          double dfmat, double *tau, double *thrv) {
  double thr = std::min(c->lk_thrcom, 1.0e-6);
  if (dfmat >= 0.0 && dfmat <= 1.0) thr *= dfmat;
  const double lkthr = std::max(thr, 1.0e-15);
  const double maxred = double(c->lk.maxred > 0 ? c->lk.maxred
                                                : int(std::max<size_t>(c->rsets.size(), 1)));
  for (int d = 0; d < nden; ++d) {
    const double no = std::max(1, nocc[d]);
    tau[d] = (lkthr / no) * dmpk;
    if (c->lk.estimate) tau[d] = tau[d] / maxred;
    thrv[d] = std::sqrt(lkthr / (no * double(c->nchol))) * dmpk;
  }
}

// Shells for the LK screening of a call: the context's (mjk_set_shells), or
// the call's own nBasSh when the context has none.
static int lk_ready(mjk_context *c, int32_t nshell, const int32_t *nbassh, const char *who) {
  if (c->lk.nshell < 1 && nshell > 0 && nbassh) {
    const int rc = mjk_set_shells(c, nshell, nbassh);
    if (rc) return rc;
  }
  if (c->lk.nshell < 1)
    return set_error(c, MJK_ERR_STATE, std::string(who) + ": LK screening needs the shells (mjk_set_shells)");
  if (!c->lk.diag_set)
    return set_error(c, MJK_ERR_STATE,
                     std::string(who) + ": LK screening needs the integral diagonal (mjk_set_lk_diag)");
  return MJK_OK;
}

// Point-group symmetry on the C1 engine: irrep s holds the global AO indices
// off[s] .. off[s+1]-1; its LT block sits at toff[s] of the symmetry-blocked
// storage (the DSBA %A0 memory).  Blocks are embedded into (and read back from)
// the full C1 lower triangle.
namespace {
struct SymMap {
  int nsym = 0;
  int64_t off[9] = {0}, toff[9] = {0};
  int64_t ntri() const { return toff[nsym]; }
  template <class F>
  void each(F &&f) const {  // f(symmetry-blocked index, full LT index)
    for (int s = 0; s < nsym; ++s) {
      const int64_t nb = off[s + 1] - off[s];
      for (int64_t a = 0; a < nb; ++a)
        for (int64_t b = 0; b <= a; ++b) {
          const int64_t ga = off[s] + a, gb = off[s] + b;
          f(toff[s] + a * (a + 1) / 2 + b, ga * (ga + 1) / 2 + gb);
        }
    }
  }
};

int sym_map(mjk_context *c, int32_t nsym, const int32_t *nbas, SymMap *m, const char *who) {
  if (nsym < 1 || nsym > 8 || !nbas)
    return set_error(c, MJK_ERR_ARG, std::string(who) + ": nsym must be in [1,8]");
  m->nsym = nsym;
  for (int s = 0; s < nsym; ++s) {
    if (nbas[s] < 0) return set_error(c, MJK_ERR_ARG, std::string(who) + ": negative nbas");
    m->off[s + 1] = m->off[s] + nbas[s];
    m->toff[s + 1] = m->toff[s] + int64_t(nbas[s]) * (nbas[s] + 1) / 2;
  }
  if (m->off[nsym] != c->n)
    return set_error(c, MJK_ERR_ARG, std::string(who) + ": sum of nbas differs from the context's");
  return MJK_OK;
}

// orbital counts per density from the (8, nden) nForb/nIorb arrays
int sym_counts(mjk_context *c, const SymMap &m, int nden, const int32_t *nforb,
               const int32_t *niorb, int32_t *nocc, const char *who) {
  for (int d = 0; d < nden; ++d) {
    nocc[d] = 0;
    for (int s = 0; s < m.nsym; ++s) {
      const int32_t f = nforb[s + 8 * d], i = niorb[s + 8 * d];
      if (f < 0 || i < 0 || f + i > m.off[s + 1] - m.off[s])
        return set_error(c, MJK_ERR_ARG, std::string(who) + ": bad orbital count");
      nocc[d] += f + i;
    }
  }
  return MJK_OK;
}
}  // namespace

namespace {
// The context's Z output, set for the duration of one call on every return path.
struct ZScope {
  Context *c;
  ZScope(Context *ctx, double *z, int32_t nocc) : c(ctx) { c->zs = ZSave{z, nocc}; }
  ~ZScope() { c->zs = ZSave{}; }
};
}  // namespace

extern "C" {

int mjk_create(mjk_context **out, int32_t nbas, int64_t nchol) {
  if (!out) return MJK_ERR_ARG;
  *out = nullptr;
  if (nbas < 1 || nbas > 46340 || nchol < 1) return MJK_ERR_ARG;
  auto *c = new (std::nothrow) mjk_context();
  if (!c) return MJK_ERR_NOMEM;
  c->n = nbas;
  c->np = int64_t(nbas) * (nbas + 1) / 2;
  c->ldl = round_up(c->np, kRowAlign);
  c->n_pad = int32_t(round_up(nbas, 128));
  c->nchol = nchol;
  c->vset.assign(size_t(nchol), 0);
  c->wsp = std::make_shared<Workspace>();
  c->w = c->wsp.get();
  int rc;
  if ((rc = hip_check(c, hipGetDevice(&c->device), "hipGetDevice")) ||
      (rc = hip_check(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking),
                      "stream create"))) {
    delete c;
    return rc;
  }
  c->own_stream = true;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess &&
      ncu > 0)
    c->ncu = ncu;
  else
    (void)hipGetLastError();
  const size_t lbytes = sizeof(double) * size_t(nchol) * size_t(c->ldl);
  void *p = nullptr;
  if (hipMalloc(&p, lbytes) != hipSuccess) {
    (void)hipGetLastError();
    hipStreamDestroy(c->stream);
    delete c;
    return MJK_ERR_NOMEM;
  }
  c->L = static_cast<double *>(p);
  if ((rc = hip_check(c, hipMemsetAsync(c->L, 0, lbytes, c->stream), "zero L")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) {
    hipFree(c->L);
    hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return MJK_OK;
}

int mjk_destroy(mjk_context *c) {
  if (!c) return MJK_OK;
  hipStreamSynchronize(c->stream);
  free_buf(c->ws_l32);  // the shared workspace goes with its last owner
  free_buf(c->lnorm);
  free_buf(c->scount);
  for (auto &r : c->rsets)
    if (r.d_map) hipFree(r.d_map);
  if (c->L) hipFree(c->L);
  for (auto e : c->ev_pool) hipEventDestroy(e);
  if (c->own_stream) hipStreamDestroy(c->stream);
  delete c;
  return MJK_OK;
}

const char *mjk_last_error(const mjk_context *c) { return c ? c->err.c_str() : "null context"; }

int mjk_set_stream(mjk_context *c, void *s) {
  if (!c) return MJK_ERR_ARG;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  if (c->own_stream) {
    (void)hipStreamDestroy(c->stream);
    c->own_stream = false;
  }
  c->stream = static_cast<hipStream_t>(s);  // NULL: the default stream
  return MJK_OK;
}

int mjk_set_ladder(mjk_context *c, int64_t j_fp32, int64_t j_bf16) {
  if (!c) return MJK_ERR_ARG;
  if (j_fp32 < 0 || j_fp32 > j_bf16 || j_bf16 > c->nchol)
    return set_error(c, MJK_ERR_ARG, "set_ladder: need 0 <= j_fp32 <= j_bf16 <= nchol");
  c->lad_fp32 = j_fp32;
  c->lad_bf16 = j_bf16;
  return MJK_OK;
}

int mjk_set_x_budget(mjk_context *c, int64_t bytes) {
  if (!c || bytes < 1) return MJK_ERR_ARG;
  c->x_budget = bytes;
  return MJK_OK;
}

int mjk_create_shared(mjk_context **out, int32_t nbas, int64_t nchol, mjk_context *share) {
  if (!out || !share) return MJK_ERR_ARG;
  int rc = mjk_create(out, nbas, nchol);
  if (rc) return rc;
  mjk_context *c = *out;
  (void)hipStreamDestroy(c->stream);
  c->stream = share->stream;  // shared workspaces need one stream
  c->own_stream = false;
  c->wsp = share->wsp;
  c->w = c->wsp.get();
  c->x_budget = share->x_budget;
  return MJK_OK;
}

int mjk_device_vectors(mjk_context *c, double **L, int64_t *ldl) {
  if (!c || !L || !ldl) return MJK_ERR_ARG;
  *L = c->L;
  *ldl = c->ldl;
  std::fill(c->vset.begin(), c->vset.end(), -1);
  c->lnorm_valid = false;
  c->nloaded = c->nchol;
  c->l32_valid = false;
  return MJK_OK;
}

int mjk_set_tuning(mjk_context *c, int32_t tile_syrk, int32_t split_target, int32_t ht_ta) {
  if (!c) return MJK_ERR_ARG;
  if ((tile_syrk != 0 && tile_syrk != 64 && tile_syrk != 128) || split_target < 0 ||
      (ht_ta != 0 && ht_ta != 64 && ht_ta != 128))
    return set_error(c, MJK_ERR_ARG, "set_tuning: tile 0/64/128, split >= 0, ht_ta 0/64/128");
  c->tile_syrk = tile_syrk;
  c->split_target = split_target;
  c->ht_ta = ht_ta;
  return MJK_OK;
}

int mjk_set_reduced_set(mjk_context *c, int32_t jred, int64_t nrs, const int64_t *map) {
  if (!c) return MJK_ERR_ARG;
  if (jred < 0 || jred > 1 << 20 || nrs < 0 || nrs > c->np || (nrs > 0 && !map))
    return set_error(c, MJK_ERR_ARG, "set_reduced_set: bad jred/nrs/map");
  std::vector<int32_t> m32(static_cast<size_t>(nrs));
  bool ident = (nrs == c->np);
  for (int64_t r = 0; r < nrs; ++r) {
    if (map[r] < 0 || map[r] >= c->np)
      return set_error(c, MJK_ERR_ARG, "set_reduced_set: LT index out of range");
    m32[size_t(r)] = int32_t(map[r]);
    ident = ident && (map[r] == r);
  }
  if (size_t(jred) >= c->rsets.size()) c->rsets.resize(size_t(jred) + 1);
  ReducedSet &rs = c->rsets[size_t(jred)];
  if (rs.d_map) {
    hipStreamSynchronize(c->stream);
    hipFree(rs.d_map);
    rs.d_map = nullptr;
  }
  rs.nrs = nrs;
  rs.identity = ident;
  rs.h_map = ident ? std::vector<int32_t>() : m32;
  if (size_t(jred) < c->lk.present_version.size()) c->lk.present_version[size_t(jred)] = -1;
  if (nrs > 0) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(int32_t) * size_t(nrs)) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(c, MJK_ERR_NOMEM, "set_reduced_set: allocation failed");
    }
    rs.d_map = static_cast<int32_t *>(p);
    int rc = hip_check(c, hipMemcpy(rs.d_map, m32.data(), sizeof(int32_t) * size_t(nrs),
                                    hipMemcpyHostToDevice),
                       "copy map");
    if (rc) return rc;
  }
  return MJK_OK;
}

int mjk_upload_vectors(mjk_context *c, int32_t jred, int64_t jvec, int64_t jnum, const double *lrs,
                       int32_t on_device) {
  if (!c) return MJK_ERR_ARG;
  if (jred < 0 || size_t(jred) >= c->rsets.size())
    return set_error(c, MJK_ERR_STATE, "upload_vectors: reduced set not declared");
  if (jvec < 0 || jnum < 0 || jvec + jnum > c->nchol)
    return set_error(c, MJK_ERR_ARG, "upload_vectors: vector range outside the context capacity");
  if (jnum == 0) return MJK_OK;
  if (!lrs) return set_error(c, MJK_ERR_ARG, "upload_vectors: null lrs");
  const ReducedSet &rs = c->rsets[size_t(jred)];
  const int64_t nrs = rs.nrs;
  int rc;
  double *dst = c->L + jvec * c->ldl;
  if (rs.identity) {
    rc = hip_check(c, hipMemcpy2DAsync(dst, sizeof(double) * c->ldl, lrs, sizeof(double) * nrs,
                                       sizeof(double) * nrs, size_t(jnum),
                                       on_device ? hipMemcpyDeviceToDevice
                                                 : hipMemcpyHostToDevice,
                                       c->stream),
                   "copy vectors");
    if (rc) return rc;
  } else {
    if ((rc = hip_check(c, hipMemsetAsync(dst, 0, sizeof(double) * c->ldl * jnum, c->stream),
                        "zero vectors")))
      return rc;
    if (nrs > 0) {
      const double *src = lrs;
      // stage host data in bounded pieces
      const int64_t per = on_device ? jnum : std::max<int64_t>(1, (int64_t(256) << 20) / (8 * nrs));
      for (int64_t j0 = 0; j0 < jnum; j0 += per) {
        const int64_t jn = std::min(per, jnum - j0);
        if (!on_device) {
          if ((rc = ensure(c, c->w->stage, sizeof(double) * nrs * jn))) return rc;
          if ((rc = hip_check(c, hipMemcpyAsync(c->w->stage.ptr, lrs + j0 * nrs,
                                                sizeof(double) * nrs * jn, hipMemcpyHostToDevice,
                                                c->stream),
                              "stage vectors")))
            return rc;
          src = static_cast<const double *>(c->w->stage.ptr);
        } else {
          src = lrs + j0 * nrs;
        }
        const unsigned gx = unsigned(std::min<int64_t>(ceil_div(nrs, 256), 1024));
        const unsigned gy = unsigned(std::min<int64_t>(jn, 65535));
        hipLaunchKernelGGL(k_scatter_rs, dim3(gx, gy), dim3(256), 0, c->stream, src, nrs, jn,
                           rs.d_map, c->L, c->ldl, jvec + j0);
        if ((rc = hip_check(c, hipGetLastError(), "scatter vectors"))) return rc;
        if (!on_device && (rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")))
          return rc;
      }
    }
  }
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "upload sync"))) return rc;
  c->l32_valid = false;
  c->lnorm_valid = false;
  for (int64_t j = jvec; j < jvec + jnum; ++j) {
    if (!c->vset[size_t(j)]) ++c->nloaded;
    c->vset[size_t(j)] = jred + 1;
  }
  return MJK_OK;
}

int mjk_upload_vectors_file(mjk_context *c, const char *path, int32_t jred, int64_t jvec,
                            int64_t jnum, const int64_t *iadr, int32_t mbl, int64_t split_bytes) {
  if (!c) return MJK_ERR_ARG;
  if (!path || (jnum > 0 && !iadr)) return set_error(c, MJK_ERR_ARG, "upload_vectors_file: null pointer");
  if (jred < 0 || size_t(jred) >= c->rsets.size())
    return set_error(c, MJK_ERR_STATE, "upload_vectors_file: reduced set not declared");
  if (jvec < 0 || jnum < 0 || jvec + jnum > c->nchol)
    return set_error(c, MJK_ERR_ARG, "upload_vectors_file: vector range outside the context capacity");
  if (split_bytes < 0) return set_error(c, MJK_ERR_ARG, "upload_vectors_file: negative split size");
  if (mbl == 0) mbl = 8;
  if (mbl != 8 && mbl != 512)
    return set_error(c, MJK_ERR_ARG, "upload_vectors_file: mbl must be 8 (word) or 512 (block)");
  if (jnum == 0) return MJK_OK;
  for (int64_t i = 0; i < jnum; ++i)
    if (iadr[i] < 0) return set_error(c, MJK_ERR_ARG, "upload_vectors_file: negative disk address");
  const int64_t nrs = c->rsets[size_t(jred)].nrs;
  // Piece k of a split unit: byte range [k*split, (k+1)*split) of the logical
  // file, stored in the base name for k = 0 and in the name with '1'..'9',
  // 'A'.. appended for k >= 1 (src/io_util/mpdafile.F90:92-140).
  constexpr int kMaxPieces = 20;  // MaxSplitFile, src/io_util/fast_io.F90:49
  struct Files {
    int fd[kMaxPieces];
    Files() { std::fill(fd, fd + kMaxPieces, -1); }
    ~Files() {
      for (int f : fd)
        if (f >= 0) close(f);
    }
  } files;
  auto open_piece = [&](int k, int *rc) -> int {
    if (k >= kMaxPieces) {
      *rc = set_error(c, MJK_ERR_SHORT_READ, "upload_vectors_file: address beyond the last file piece");
      return -1;
    }
    if (files.fd[k] < 0) {
      std::string name(path);
      if (k > 0) name += char(k <= 9 ? '0' + k : 'A' + (k - 10));
      files.fd[k] = open(name.c_str(), O_RDONLY);
      if (files.fd[k] < 0) {
        *rc = set_error(c, MJK_ERR_IO, "upload_vectors_file: cannot open " + name + ": " +
                                           std::strerror(errno));
        return -1;
      }
    }
    return files.fd[k];
  };
  // batches of at most 256 MiB through one host buffer, each uploaded like Lrs
  const int64_t per = std::max<int64_t>(1, (int64_t(256) << 20) / (8 * std::max<int64_t>(nrs, 1)));
  std::vector<double> buf(size_t(std::min(per, jnum) * nrs));
  for (int64_t j0 = 0; j0 < jnum; j0 += per) {
    const int64_t jn = std::min(per, jnum - j0);
    for (int64_t i = 0; i < jn; ++i) {
      char *dst = reinterpret_cast<char *>(buf.data() + i * nrs);
      int64_t left = nrs * int64_t(sizeof(double));
      int64_t at = iadr[j0 + i] * int64_t(mbl);  // disk address in units of mbl bytes
      while (left > 0) {
        const int k = split_bytes > 0 ? int(at / split_bytes) : 0;
        const int64_t in = split_bytes > 0 ? at - k * split_bytes : at;
        const int64_t room = split_bytes > 0 ? std::min(left, split_bytes - in) : left;
        int rc = MJK_OK;
        const int fd = open_piece(k, &rc);
        if (fd < 0) return rc;
        const ssize_t got = pread(fd, dst, size_t(room), off_t(in));
        if (got < 0) {
          if (errno == EINTR) continue;
          return set_error(c, MJK_ERR_IO, std::string("upload_vectors_file: read failed: ") +
                                              std::strerror(errno));
        }
        if (got == 0)
          return set_error(c, MJK_ERR_SHORT_READ, "upload_vectors_file: file ends inside a vector");
        dst += got;
        left -= got;
        at += got;
      }
    }
    const int rc = mjk_upload_vectors(c, jred, jvec + j0, jn, buf.data(), 0);
    if (rc) return rc;
  }
  return MJK_OK;
}

int mjk_coulomb(mjk_context *c, int32_t nd, const double *dlt, int64_t ldd, double *jlt,
                int64_t ldj, double *v_out) {
  if (!c) return MJK_ERR_ARG;
  if (!dlt || !jlt) return set_error(c, MJK_ERR_ARG, "coulomb: null pointer");
  return coulomb(c, nd, dlt, ldd, jlt, ldj, v_out);
}

int mjk_coulomb_noalloc(mjk_context *c, int32_t nd, const double *dlt, int64_t ldd, double *jlt,
                        int64_t ldj, double *v_out) {
  if (!c) return MJK_ERR_ARG;
  if (!dlt || !jlt) return set_error(c, MJK_ERR_ARG, "coulomb_noalloc: null pointer");
  return coulomb_noalloc(c, nd, dlt, ldd, jlt, ldj, v_out);
}

int mjk_exchange(mjk_context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                 double *klt, int64_t ldk, int32_t precision) {
  if (!c) return MJK_ERR_ARG;
  if (!nocc || !corb || !klt) return set_error(c, MJK_ERR_ARG, "exchange: null pointer");
  for (int d = 0; d < nk; ++d)
    if (nocc[d] > 0 && !corb[d]) return set_error(c, MJK_ERR_ARG, "exchange: null orbital matrix");
  return exchange(c, nk, nocc, corb, klt, ldk, precision);
}

int mjk_coulomb_exchange(mjk_context *c, const double *dlt, int64_t ldd, double *jlt, int64_t ldj,
                         int32_t nk, const int32_t *nocc, const double *const *corb, double *klt,
                         int64_t ldk, double *v_out) {
  if (!c) return MJK_ERR_ARG;
  if (!dlt || !jlt || !nocc || !corb || !klt)
    return set_error(c, MJK_ERR_ARG, "coulomb_exchange: null pointer");
  if (nk < 1 || nk > 16) return set_error(c, MJK_ERR_ARG, "exchange: nk must be in [1,16]");
  for (int d = 0; d < nk; ++d)
    if (nocc[d] > 0 && !corb[d]) return set_error(c, MJK_ERR_ARG, "exchange: null orbital matrix");
  return coulomb_exchange(c, dlt, ldd, jlt, ldj, nk, nocc, corb, klt, ldk, v_out);
}

int mjk_exchange_z(mjk_context *c, int32_t nocc, const double *corb, double *klt,
                   int32_t precision, double *z) {
  if (!c) return MJK_ERR_ARG;
  if (!corb || !klt || !z) return set_error(c, MJK_ERR_ARG, "exchange_z: null pointer");
  if (nocc < 1 || nocc > 256) return set_error(c, MJK_ERR_ARG, "exchange_z: nocc must be in [1,256]");
  ZScope scope(c, z, nocc);
  const double *sets[1] = {corb};
  return exchange(c, 1, &nocc, sets, klt, c->np, precision);
}

int mjk_coulomb_exchange_z(mjk_context *c, const double *dlt, double *jlt, int32_t nocc,
                           const double *corb, double *klt, double *v_out, double *z) {
  if (!c) return MJK_ERR_ARG;
  if (!dlt || !jlt || !corb || !klt || !z)
    return set_error(c, MJK_ERR_ARG, "coulomb_exchange_z: null pointer");
  if (nocc < 1 || nocc > 256)
    return set_error(c, MJK_ERR_ARG, "coulomb_exchange_z: nocc must be in [1,256]");
  ZScope scope(c, z, nocc);
  const double *sets[1] = {corb};
  return coulomb_exchange(c, dlt, c->np, jlt, c->np, 1, &nocc, sets, klt, c->np, v_out);
}

int mjk_exchange_mixed(mjk_context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                       double *klt, int64_t ldk, int64_t j_fp32, int64_t j_bf16) {
  if (!c) return MJK_ERR_ARG;
  if (!nocc || !corb || !klt) return set_error(c, MJK_ERR_ARG, "exchange: null pointer");
  for (int d = 0; d < nk; ++d)
    if (nocc[d] > 0 && !corb[d]) return set_error(c, MJK_ERR_ARG, "exchange: null orbital matrix");
  return exchange_mixed(c, nk, nocc, corb, klt, ldk, j_fp32, j_bf16);
}

int mjk_vector_norms(mjk_context *c, double *w_dev) {
  if (!c) return MJK_ERR_ARG;
  return vector_norms(c, w_dev);
}

int mjk_exchange_screened(mjk_context *c, int32_t nk, const int32_t *nocc,
                          const double *const *corb, double *klt, int64_t ldk, int32_t precision,
                          double tau, double thr_x) {
  if (!c) return MJK_ERR_ARG;
  if (!nocc || !corb || !klt) return set_error(c, MJK_ERR_ARG, "exchange: null pointer");
  for (int d = 0; d < nk; ++d)
    if (nocc[d] > 0 && !corb[d]) return set_error(c, MJK_ERR_ARG, "exchange: null orbital matrix");
  return exchange_screened(c, nk, nocc, corb, klt, ldk, precision, tau, thr_x);
}

int mjk_screening_stats(mjk_context *c, int64_t *k_items, int64_t *k_skipped, int64_t *x_tiles,
                        int64_t *x_skipped) {
  if (!c) return MJK_ERR_ARG;
  unsigned long long h[2] = {0, 0};
  if (c->scount.ptr) {
    int rc;
    if ((rc = hip_check(c, hipMemcpyAsync(h, c->scount.ptr, sizeof(h), hipMemcpyDeviceToHost,
                                          c->stream), "screening counters")) ||
        (rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")))
      return rc;
  }
  if (k_items) *k_items = c->k_items;
  if (k_skipped) *k_skipped = int64_t(h[0]);
  if (x_tiles) *x_tiles = c->x_tiles;
  if (x_skipped) *x_skipped = int64_t(h[1]);
  return MJK_OK;
}

int mjk_set_lk_screening(mjk_context *c, int32_t enable, double thrcom) {
  if (!c) return MJK_ERR_ARG;
  if (enable && !(thrcom > 0.0)) return set_error(c, MJK_ERR_ARG, "set_lk_screening: thrcom <= 0");
  c->lk_on = enable != 0;
  if (enable) c->lk_thrcom = thrcom;
  return MJK_OK;
}

int mjk_set_shells(mjk_context *c, int32_t nshell, const int32_t *nbassh) {
  if (!c) return MJK_ERR_ARG;
  if (nshell < 1 || !nbassh) return set_error(c, MJK_ERR_ARG, "set_shells: need nshell >= 1 shells");
  int64_t sum = 0;
  for (int s = 0; s < nshell; ++s) {
    if (nbassh[s] < 0) return set_error(c, MJK_ERR_ARG, "set_shells: negative shell size");
    sum += nbassh[s];
  }
  if (sum != c->n) return set_error(c, MJK_ERR_ARG, "set_shells: shell sizes do not add up to nbas");
  LKState &K = c->lk;
  std::vector<int32_t> shell(size_t(c->n)), koff(size_t(nshell) + 1);
  for (int s = 0, a = 0; s < nshell; ++s) {
    koff[size_t(s)] = a;
    for (int i = 0; i < nbassh[s]; ++i) shell[size_t(a++)] = s;
  }
  koff[size_t(nshell)] = c->n;
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")) ||
      (rc = ensure(c, K.shell, sizeof(int32_t) * size_t(c->n))) ||
      (rc = ensure(c, K.koff, sizeof(int32_t) * (size_t(nshell) + 1))) ||
      (rc = ensure(c, K.nbsh, sizeof(int32_t) * size_t(nshell))) ||
      (rc = hip_check(c, hipMemcpy(K.shell.ptr, shell.data(), sizeof(int32_t) * size_t(c->n),
                                   hipMemcpyHostToDevice), "copy shells")) ||
      (rc = hip_check(c, hipMemcpy(K.koff.ptr, koff.data(), sizeof(int32_t) * (size_t(nshell) + 1),
                                   hipMemcpyHostToDevice), "copy shell offsets")) ||
      (rc = hip_check(c, hipMemcpy(K.nbsh.ptr, nbassh, sizeof(int32_t) * size_t(nshell),
                                   hipMemcpyHostToDevice), "copy shell sizes")))
    return rc;
  K.nshell = nshell;
  K.nbassh.assign(nbassh, nbassh + nshell);
  K.version += 1;  // shell-pair occupancy of every reduced set is stale
  return MJK_OK;
}

int mjk_set_lk_diag(mjk_context *c, const double *diag, int32_t maxred, int32_t estimate,
                    int32_t update) {
  if (!c) return MJK_ERR_ARG;
  if (!diag) return set_error(c, MJK_ERR_ARG, "set_lk_diag: null diagonal");
  if (maxred < 0) return set_error(c, MJK_ERR_ARG, "set_lk_diag: maxred < 0");
  LKState &K = c->lk;
  int rc;
  if ((rc = ensure(c, K.diag0, sizeof(double) * size_t(c->np))) ||
      (rc = hip_check(c, hipMemcpy(K.diag0.ptr, diag, sizeof(double) * size_t(c->np),
                                   hipMemcpyHostToDevice), "copy diagonal")))
    return rc;
  K.diag_set = true;
  K.maxred = maxred;
  K.estimate = estimate != 0;
  K.update = update != 0;
  return MJK_OK;
}

int mjk_lk_stats(mjk_context *c, int64_t *stats) {
  if (!c) return MJK_ERR_ARG;
  if (!stats) return set_error(c, MJK_ERR_ARG, "lk_stats: null pointer");
  for (int i = 0; i < 3; ++i) stats[i] = int64_t(c->lk.last_stats[i]);
  stats[3] = c->lk.last_batches;
  return MJK_OK;
}

int mjk_fock_update(mjk_context *c, int32_t nd, double *flt, int64_t ldf, const double *jlt,
                    int64_t ldj, const int32_t *jsrc, const double *fj, const double *klt,
                    int64_t ldk, const double *fk) {
  if (!c) return MJK_ERR_ARG;
  if (!flt) return set_error(c, MJK_ERR_ARG, "fock_update: null flt");
  return fock_update(c, nd, flt, ldf, jlt, ldj, jsrc, fj, klt, ldk, fk);
}

int mjk_energy(mjk_context *c, const double *h, const double *dlt, const double *g, double *e) {
  if (!c) return MJK_ERR_ARG;
  return energy(c, h, dlt, g, e);
}

int mjk_sync(mjk_context *c) {
  if (!c) return MJK_ERR_ARG;
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

int mjk_set_profiling(mjk_context *c, int32_t on) {
  if (!c) return MJK_ERR_ARG;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  c->pending.clear();
  c->ev_used = 0;
  c->ev_open = -1;
  c->profile = on != 0;
  for (int i = 0; i < kNumPhases; ++i) {
    c->phase_ms[i] = 0.0;
    c->phase_calls[i] = 0;
  }
  return MJK_OK;
}

int mjk_phase_times(mjk_context *c, double *ms, int64_t *calls) {
  if (!c) return MJK_ERR_ARG;
  int rc = resolve_phases(c);
  if (rc) return rc;
  for (int i = 0; i < kNumPhases; ++i) {
    if (ms) ms[i] = c->phase_ms[i];
    if (calls) calls[i] = c->phase_calls[i];
  }
  return MJK_OK;
}

// ---- reference-shaped drop-ins -------------------------------------------

int mjk_cho_fscf(mjk_context *c, int32_t nden, double *const *flt, const int32_t *nforb,
                 const int32_t *niorb, const double *const *porb, const double *dlt, double exfac,
                 int32_t precision) {
  if (!c) return MJK_ERR_ARG;
  if (nden != 1 && nden != 2)
    return set_error(c, MJK_ERR_ARG, "cho_fscf: invalid parameter nDen");  // cho_fscf.F90:75-78
  if (!flt || !nforb || !niorb || !dlt) return set_error(c, MJK_ERR_ARG, "cho_fscf: null pointer");
  const int64_t np = c->np;
  int32_t nocc[2] = {0, 0};
  int64_t tot = 0;
  for (int d = 0; d < nden; ++d) {
    if (nforb[d] < 0 || niorb[d] < 0) return set_error(c, MJK_ERR_ARG, "cho_fscf: negative nOcc");
    if (!flt[d]) return set_error(c, MJK_ERR_ARG, "cho_fscf: null FLT(d)");
    nocc[d] = nforb[d] + niorb[d];
    if (nocc[d] > 0 && (!porb || !porb[d]))
      return set_error(c, MJK_ERR_ARG, "cho_fscf: null Porb(d)");
    tot += int64_t(nocc[d]) * c->n;
  }
  // device scratch: D | J | K(nden) | F(nden) | orbitals
  const size_t need = sizeof(double) * size_t(np * (2 + 2 * nden) + tot);
  int rc;
  if ((rc = ensure(c, c->w->misc, need))) return rc;
  double *dD = static_cast<double *>(c->w->misc.ptr);
  double *dJ = dD + np, *dK = dJ + np, *dF = dK + nden * np, *dC = dF + nden * np;
  const double *cptr[2] = {dC, dC + int64_t(nocc[0]) * c->n};
  if ((rc = hip_check(c, hipMemcpyAsync(dD, dlt, sizeof(double) * np, hipMemcpyHostToDevice,
                                        c->stream), "copy DLT")))
    return rc;
  for (int d = 0; d < nden; ++d) {  // FLT(d) and Porb(d) are separate arrays
    if ((rc = hip_check(c, hipMemcpyAsync(dF + d * np, flt[d], sizeof(double) * np,
                                          hipMemcpyHostToDevice, c->stream), "copy FLT")))
      return rc;
    if (nocc[d] > 0 &&
        (rc = hip_check(c, hipMemcpyAsync(const_cast<double *>(cptr[d]), porb[d],
                                          sizeof(double) * size_t(nocc[d]) * c->n,
                                          hipMemcpyHostToDevice, c->stream), "copy Porb")))
      return rc;
  }
  if (precision == MJK_PREC_FP64) {  // V pass inside the half-transform
    if ((rc = coulomb_exchange(c, dD, np, dJ, np, nden, nocc, cptr, dK, np, nullptr))) return rc;
  } else {
    if ((rc = coulomb(c, 1, dD, np, dJ, np, nullptr))) return rc;
    if ((rc = exchange(c, nden, nocc, cptr, dK, np, precision))) return rc;
  }
  const int32_t jsrc[2] = {0, 0};
  const double fj[2] = {1.0, 1.0};                 // FactCI, cho_fscf.F90:64
  const double fk[2] = {-exfac, -exfac};           // FactXI, cho_fscf.F90:65
  if ((rc = fock_update(c, nden, dF, np, dJ, np, jsrc, fj, dK, np, fk))) return rc;
  for (int d = 0; d < nden; ++d)
    if ((rc = hip_check(c, hipMemcpyAsync(flt[d], dF + d * np, sizeof(double) * np,
                                          hipMemcpyDeviceToHost, c->stream), "copy FLT back")))
      return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "cho_fscf sync");
}

int mjk_cho_fscf_sym(mjk_context *c, int32_t nsym, const int32_t *nbas, int32_t nden,
                     double *const *flt, const int32_t *nforb, const int32_t *niorb,
                     const double *const *porb, const double *dlt, double exfac,
                     int32_t precision) {
  if (!c) return MJK_ERR_ARG;
  if (nden != 1 && nden != 2)
    return set_error(c, MJK_ERR_ARG, "cho_fscf_sym: invalid parameter nDen");  // cho_fscf.F90:75-78
  if (!flt || !nforb || !niorb || !dlt) return set_error(c, MJK_ERR_ARG, "cho_fscf_sym: null pointer");
  SymMap m;
  int32_t nocc[2] = {0, 0};
  int rc;
  if ((rc = sym_map(c, nsym, nbas, &m, "cho_fscf_sym")) ||
      (rc = sym_counts(c, m, nden, nforb, niorb, nocc, "cho_fscf_sym")))
    return rc;
  const int64_t n = c->n, np = c->np;
  for (int d = 0; d < nden; ++d)
    if (!flt[d] || (nocc[d] > 0 && (!porb || !porb[d])))
      return set_error(c, MJK_ERR_ARG, "cho_fscf_sym: null FLT(d) or Porb(d)");
  // DLT and FLT blocks into the full LT; the orbitals of irrep s into columns of
  // orbital-major A2(nocc_d, n) matrices that vanish outside the irrep
  std::vector<double> dfull(static_cast<size_t>(np), 0.0), ffull(static_cast<size_t>(np * nden), 0.0);
  std::vector<double> cfull(static_cast<size_t>((int64_t(nocc[0]) + nocc[1]) * n), 0.0);
  m.each([&](int64_t t, int64_t g) {
    dfull[size_t(g)] = dlt[t];
    for (int d = 0; d < nden; ++d) ffull[size_t(d * np + g)] = flt[d][t];
  });
  double *cd = cfull.data();
  for (int d = 0; d < nden; ++d) {
    const double *pd = porb ? porb[d] : nullptr;
    int64_t k0 = 0;  // first orbital of irrep s within density d
    for (int s = 0; s < nsym; ++s) {
      const int64_t no = nforb[s + 8 * d] + niorb[s + 8 * d], nb = nbas[s];
      for (int64_t b = 0; b < nb; ++b)
        for (int64_t k = 0; k < no; ++k) cd[(k0 + k) + nocc[d] * (m.off[s] + b)] = pd[k + no * b];
      pd += no * nb;
      k0 += no;
    }
    cd += int64_t(nocc[d]) * n;
  }
  const int32_t nf0[2] = {0, 0};
  double *fp[2] = {ffull.data(), ffull.data() + np};
  const double *cp[2] = {cfull.data(), cfull.data() + int64_t(nocc[0]) * n};
  if ((rc = mjk_cho_fscf(c, nden, fp, nf0, nocc, cp, dfull.data(), exfac, precision))) return rc;
  m.each([&](int64_t t, int64_t g) {
    for (int d = 0; d < nden; ++d) flt[d][t] = ffull[size_t(d * np + g)];
  });
  return MJK_OK;
}

int mjk_cho_lk_scf_sym(mjk_context *c, int32_t nsym, const int32_t *nbas, int32_t nden,
                       double *const *flt, double *const *klt, const int32_t *nforb,
                       const int32_t *niorb, const double *const *porb, const double *plt,
                       double factxi, int32_t nscreen, double dmpk, double dfmat,
                       int32_t nshell, const int32_t *nbassh, int32_t precision) {
  if (!c) return MJK_ERR_ARG;
  if (nden != 1 && nden != 2)
    return set_error(c, MJK_ERR_ARG, "cho_lk_scf_sym: invalid parameter nDen");  // cho_lk_scf.F90:93-96
  if (!flt || !nforb || !niorb || !plt) return set_error(c, MJK_ERR_ARG, "cho_lk_scf_sym: null pointer");
  if (c->lk_on && nsym > 1)  // the reference screens per irrep shell blocks (nBasSh(iSym,iSh))
    return set_error(c, MJK_ERR_ARG, "cho_lk_scf_sym: LK screening with nSym > 1 is not supported");
  SymMap m;
  int32_t nocc[2] = {0, 0};
  int rc;
  if ((rc = sym_map(c, nsym, nbas, &m, "cho_lk_scf_sym")) ||
      (rc = sym_counts(c, m, nden, nforb, niorb, nocc, "cho_lk_scf_sym")))
    return rc;
  const int64_t n = c->n, np = c->np;
  for (int d = 0; d < nden; ++d)
    if (!flt[d] || (nocc[d] > 0 && (!porb || !porb[d])) || (klt && !klt[d]))
      return set_error(c, MJK_ERR_ARG, "cho_lk_scf_sym: null FLT(d), KLT(d) or Porb(d)");
  // KLT(d)%SB(s): shell-pair blocks of irrep s's shells (cho_lk_scf.F90:274-292)
  std::vector<int64_t> kpos;  // full LT index of the embedded basis -> KLT(d)%A0 index
  if (klt) {
    kpos.assign(size_t(np), -1);
    for (int s = 0; s < nsym; ++s) {
      std::vector<int64_t> sh;
      int64_t sum = 0;
      for (int i = 0; i < (nshell > 0 && nbassh ? nshell : 1); ++i) {
        const int64_t v = nshell > 0 && nbassh ? nbassh[s + nsym * i] : nbas[s];
        if (v < 0) return set_error(c, MJK_ERR_ARG, "cho_lk_scf_sym: negative shell size");
        sh.push_back(v);
        sum += v;
      }
      if (sum != nbas[s])
        return set_error(c, MJK_ERR_ARG, "cho_lk_scf_sym: shell sizes of an irrep do not add up to its nbas");
      int64_t off = m.toff[s], first_a = 0;
      for (size_t sa = 0; sa < sh.size(); ++sa) {
        int64_t first_b = 0;
        for (size_t sb = 0; sb <= sa; ++sb) {
          for (int64_t bl = 0; bl < sh[sb]; ++bl)
            for (int64_t al = (sa == sb ? bl : 0); al < sh[sa]; ++al) {
              const int64_t a = m.off[s] + first_a + al, b = m.off[s] + first_b + bl;
              kpos[size_t(a * (a + 1) / 2 + b)] = off + (sa == sb ? al * (al + 1) / 2 + bl : sh[sa] * bl + al);
            }
          off += sa == sb ? sh[sa] * (sh[sa] + 1) / 2 : sh[sa] * sh[sb];
          first_b += sh[sb];
        }
        first_a += sh[sa];
      }
    }
  }
  // Porb(d)%SB(s)%A2(nBas(s),nBas(s)), AO index fastest, its first nForb+nIorb
  // columns the orbitals: they become columns k0(s).. of an n x n matrix
  std::vector<double> pfull(static_cast<size_t>(np), 0.0), ffull(static_cast<size_t>(np * nden), 0.0);
  std::vector<double> cfull(static_cast<size_t>(nden * n * n), 0.0);
  m.each([&](int64_t t, int64_t g) {
    pfull[size_t(g)] = plt[t];
    for (int d = 0; d < nden; ++d) ffull[size_t(d * np + g)] = flt[d][t];
  });
  for (int d = 0; d < nden; ++d) {
    const double *pd = porb ? porb[d] : nullptr;
    double *cd = cfull.data() + d * n * n;
    int64_t k0 = 0;
    for (int s = 0; s < nsym; ++s) {
      const int64_t no = nforb[s + 8 * d] + niorb[s + 8 * d], nb = nbas[s];
      for (int64_t k = 0; k < no; ++k)
        for (int64_t b = 0; b < nb; ++b) cd[(m.off[s] + b) + n * (k0 + k)] = pd[b + nb * k];
      pd += nb * nb;
      k0 += no;
    }
  }
  const int32_t nf0[2] = {0, 0};
  double *fp[2] = {ffull.data(), ffull.data() + np};
  const double *cp[2] = {cfull.data(), cfull.data() + n * n};
  std::vector<double> kfull(klt ? size_t(np * nden) : 0, 0.0);  // -FactXI K, plain LT of the embedding
  double *kp[2] = {kfull.data(), kfull.data() + (klt ? np : 0)};
  if ((rc = mjk_cho_lk_scf(c, nden, fp, klt ? kp : nullptr, nf0, nocc, cp, pfull.data(), factxi, nscreen,
                           dmpk, dfmat, 0, nullptr, precision)))
    return rc;
  m.each([&](int64_t t, int64_t g) {
    for (int d = 0; d < nden; ++d) flt[d][t] = ffull[size_t(d * np + g)];
  });
  if (klt)  // K is block diagonal: only the irrep blocks have a place in KLT
    for (int d = 0; d < nden; ++d)
      for (int64_t g = 0; g < np; ++g)
        if (kpos[size_t(g)] >= 0) klt[d][kpos[size_t(g)]] += kfull[size_t(d * np + g)];
  return MJK_OK;
}

int mjk_cho_lk_scf(mjk_context *c, int32_t nden, double *const *flt, double *const *klt,
                   const int32_t *nforb, const int32_t *niorb, const double *const *porb,
                   const double *plt, double factxi,
                   int32_t nscreen, double dmpk, double dfmat, int32_t nshell,
                   const int32_t *nbassh, int32_t precision) {
  // Without mjk_set_lk_screening every shell pair and orbital is computed (the
  // reference's result for dmpk -> 0); with it, the reference's LK screening.
  if (!c) return MJK_ERR_ARG;
  if (nden != 1 && nden != 2)
    return set_error(c, MJK_ERR_ARG, "cho_lk_scf: invalid parameter nDen");  // cho_lk_scf.F90:93-96
  if (!flt || !nforb || !niorb || !plt) return set_error(c, MJK_ERR_ARG, "cho_lk_scf: null pointer");
  if (dmpk < 0.0) return set_error(c, MJK_ERR_ARG, "cho_lk_scf: dmpk < 0");
  const int64_t np = c->np, n = c->n;
  int rc0;
  int32_t nocc[2] = {0, 0};
  int64_t tot = 0;
  for (int d = 0; d < nden; ++d) {
    if (nforb[d] < 0 || niorb[d] < 0 || nforb[d] + niorb[d] > n)
      return set_error(c, MJK_ERR_ARG, "cho_lk_scf: bad orbital count");
    nocc[d] = nforb[d] + niorb[d];
    tot += int64_t(nocc[d]) * n;
    if (!flt[d] || (nocc[d] > 0 && (!porb || !porb[d])) || (klt && !klt[d]))
      return set_error(c, MJK_ERR_ARG, "cho_lk_scf: null FLT(d), KLT(d) or Porb(d)");
  }
  // KLT storage: shell-pair blocks (cho_lk_scf.F90:274-292); no shells = one shell
  std::vector<int64_t> kpos;
  if (klt) {
    std::vector<int32_t> sh;
    if (nshell > 0 && nbassh) sh.assign(nbassh, nbassh + nshell);
    else sh.assign(1, int32_t(n));
    int64_t sum = 0;
    for (int32_t v : sh) {
      if (v < 0) return set_error(c, MJK_ERR_ARG, "cho_lk_scf: negative shell size");
      sum += v;
    }
    if (sum != n) return set_error(c, MJK_ERR_ARG, "cho_lk_scf: shell sizes do not add up to nbas");
    kpos.resize(static_cast<size_t>(np));
    int64_t off = 0, first_a = 0;
    for (size_t sa = 0; sa < sh.size(); ++sa) {
      int64_t first_b = 0;
      for (size_t sb = 0; sb <= sa; ++sb) {
        for (int64_t bl = 0; bl < sh[sb]; ++bl)
          for (int64_t al = (sa == sb ? bl : 0); al < sh[sa]; ++al) {
            const int64_t a = first_a + al, b = first_b + bl;
            kpos[size_t(a * (a + 1) / 2 + b)] =
                off + (sa == sb ? al * (al + 1) / 2 + bl : sh[sa] * bl + al);
          }
        off += sa == sb ? int64_t(sh[sa]) * (sh[sa] + 1) / 2 : int64_t(sh[sa]) * sh[sb];
        first_b += sh[sb];
      }
      first_a += sh[sa];
    }
  }
  if (c->lk_on && (rc0 = lk_ready(c, nshell, nbassh, "cho_lk_scf"))) return rc0;
  // orbitals to the orbital-major layout the exchange engine reads
  std::vector<double> corb(static_cast<size_t>(tot));
  for (int64_t d = 0, o0 = 0; d < nden; o0 += int64_t(nocc[d]) * n, ++d)
    for (int64_t k = 0; k < nocc[d]; ++k)
      for (int64_t b = 0; b < n; ++b)
        corb[size_t(o0 + k + nocc[d] * b)] = porb[d][b + n * k];
  // device scratch: D | J | K(nden) | F(nden) | orbitals
  const size_t need = sizeof(double) * size_t(np * (2 + 2 * nden) + tot);
  int rc;
  if ((rc = ensure(c, c->w->misc, need))) return rc;
  double *dD = static_cast<double *>(c->w->misc.ptr);
  double *dJ = dD + np, *dK = dJ + np, *dF = dK + nden * np, *dC = dF + nden * np;
  if ((rc = hip_check(c, hipMemcpyAsync(dD, plt, sizeof(double) * np, hipMemcpyHostToDevice,
                                        c->stream), "copy PLT")))
    return rc;
  for (int d = 0; d < nden; ++d)
    if ((rc = hip_check(c, hipMemcpyAsync(dF + d * np, flt[d], sizeof(double) * np,
                                          hipMemcpyHostToDevice, c->stream), "copy FLT")))
      return rc;
  if (tot > 0 && (rc = hip_check(c, hipMemcpyAsync(dC, corb.data(), sizeof(double) * tot,
                                                   hipMemcpyHostToDevice, c->stream),
                                 "copy Porb")))
    return rc;
  const double *cptr[2] = {dC, dC + int64_t(nocc[0]) * n};
  if (c->lk_on) {  // the reference's LK screening, FP64
    if (precision != MJK_PREC_FP64)
      return set_error(c, MJK_ERR_ARG, "cho_lk_scf: LK screening computes in FP64 only");
    double tau[2], thrv[2];
    lk_thresholds(c, nden, nocc, dmpk, dfmat, tau, thrv);
    c->lk_nscreen = nscreen;
    if ((rc = coulomb(c, 1, dD, np, dJ, np, nullptr)) ||
        (rc = lk_exchange(c, nden, nocc, cptr, tau, thrv, false, dK, np)))
      return rc;
  } else if (precision == MJK_PREC_FP64) {  // V pass inside the half-transform
    if ((rc = coulomb_exchange(c, dD, np, dJ, np, nden, nocc, cptr, dK, np, nullptr))) return rc;
  } else {
    if ((rc = coulomb(c, 1, dD, np, dJ, np, nullptr))) return rc;
    if ((rc = exchange(c, nden, nocc, cptr, dK, np, precision))) return rc;
  }
  const int32_t jsrc[2] = {0, 0};
  const double fj[2] = {1.0, 1.0};               // FactCI, cho_lk_scf.F90:102
  const double fk[2] = {-factxi, -factxi};       // -FactXI, cho_lk_scf.F90:803,830
  if ((rc = fock_update(c, nden, dF, np, dJ, np, jsrc, fj, dK, np, fk))) return rc;
  for (int d = 0; d < nden; ++d)
    if ((rc = hip_check(c, hipMemcpyAsync(flt[d], dF + d * np, sizeof(double) * np,
                                          hipMemcpyDeviceToHost, c->stream), "copy FLT back")))
      return rc;
  std::vector<double> kh;
  if (klt) {
    kh.resize(static_cast<size_t>(np * nden));
    if ((rc = hip_check(c, hipMemcpyAsync(kh.data(), dK, sizeof(double) * np * nden,
                                          hipMemcpyDeviceToHost, c->stream), "copy K back")))
      return rc;
  }
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "cho_lk_scf sync"))) return rc;
  if (klt)
    for (int d = 0; d < nden; ++d)
      for (int64_t p = 0; p < np; ++p)
        klt[d][kpos[size_t(p)]] += -factxi * kh[size_t(d * np + p)];
  return MJK_OK;
}

int mjk_cho_fock_dft_red(mjk_context *c, const double *dlt, double *flt) {
  if (!c) return MJK_ERR_ARG;
  if (!dlt || !flt) return set_error(c, MJK_ERR_ARG, "cho_fock_dft_red: null pointer");
  const int64_t np = c->np;
  int rc;
  if ((rc = ensure(c, c->w->misc, sizeof(double) * size_t(2 * np)))) return rc;
  double *dD = static_cast<double *>(c->w->misc.ptr);
  double *dJ = dD + np;
  if ((rc = hip_check(c, hipMemcpyAsync(dD, dlt, sizeof(double) * np, hipMemcpyHostToDevice,
                                        c->stream), "copy DLT")))
    return rc;
  if ((rc = coulomb(c, 1, dD, np, dJ, np, nullptr))) return rc;
  // FLT is written, not accumulated: the first reduced set's swap_rs2full
  // (add = .false., cho_fock_dft_red.F90:190-191) zeroes all of FLT
  // (swap_rs2full.F90:37-41) and later sets add to it.
  if ((rc = hip_check(c, hipMemcpyAsync(flt, dJ, sizeof(double) * np, hipMemcpyDeviceToHost,
                                        c->stream), "copy FLT back")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "cho_fock_dft_red sync");
}

}  // extern "C"

// FI += J[DI] + fk_i * K[porb1], FA += J[DA] + fk_a * K[porb2] (if do_active),
// W_PWXY from porb3 and cmo: the shared body of the CASSCF drop-ins.
// porb1/porb2/porb3 are host arrays in the orbital-major layout A2(nk,nBas).
static int fmcscf_core(mjk_context *c, double *const *flt, int32_t nforb, int32_t niorb,
                       int32_t naorb, double fk_i, double fk_a, const double *const *dlt,
                       int32_t do_active,
                       const double *porb1, int32_t nchm, const double *porb2,
                       const double *porb3, double *w_pwxy, const double *cmo, int32_t nfro,
                       int32_t norb, int32_t precision, const double *tau = nullptr,
                       const double *thrv = nullptr) {
  if (!flt || !dlt) return set_error(c, MJK_ERR_ARG, "cho_fmcscf: null pointer");
  if (nforb < 0 || niorb < 0 || naorb < 0 || nchm < 0)
    return set_error(c, MJK_ERR_ARG, "cho_fmcscf: negative orbital count");
  const int nden = do_active ? 2 : 1;
  for (int d = 0; d < nden; ++d)
    if (!flt[d] || !dlt[d]) return set_error(c, MJK_ERR_ARG, "cho_fmcscf: null FLT(d) or DLT(d)");
  const int64_t np = c->np, n = c->n;
  const int32_t nk1 = nforb + niorb;
  const int32_t nk2 = do_active ? nchm : 0;
  if ((nk1 > 0 && !porb1) || (nk2 > 0 && !porb2))
    return set_error(c, MJK_ERR_ARG, "cho_fmcscf: null orbital matrix");
  const bool do_puvx = naorb > 0 && w_pwxy;
  if (do_puvx && (!porb3 || !cmo || nfro < 0 || norb < 0 || nfro + norb > n))
    return set_error(c, MJK_ERR_ARG, "cho_fmcscf: bad (pw|xy) arguments");
  const int64_t nxy = int64_t(naorb) * (naorb + 1) / 2;
  const int64_t nw = do_puvx ? int64_t(norb) * naorb * nxy : 0;
  const size_t need = sizeof(double) * size_t(np * 4 * nden + (nk1 + nk2) * n +
                                              (do_puvx ? naorb * n + n * n + nw : 0));
  int rc;
  if ((rc = ensure(c, c->w->misc, need))) return rc;
  double *dD = static_cast<double *>(c->w->misc.ptr);
  double *dJ = dD + nden * np, *dK = dJ + nden * np, *dF = dK + nden * np, *dC1 = dF + nden * np;
  double *dC2 = dC1 + int64_t(nk1) * n, *dC3 = dC2 + int64_t(nk2) * n;
  double *dCMO = dC3 + (do_puvx ? int64_t(naorb) * n : 0), *dW = dCMO + (do_puvx ? n * n : 0);
  auto h2d = [&](void *dst, const void *src, size_t cnt, const char *what) {
    return hip_check(c, hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice,
                                       c->stream), what);
  };
  for (int d = 0; d < nden; ++d)  // FI/FA and DI/DA are separate arrays
    if ((rc = h2d(dD + d * np, dlt[d], size_t(np), "copy DLT")) ||
        (rc = h2d(dF + d * np, flt[d], size_t(np), "copy FLT")))
      return rc;
  if (nk1 > 0 && (rc = h2d(dC1, porb1, size_t(nk1 * n), "copy POrb(1)"))) return rc;
  if (nk2 > 0 && (rc = h2d(dC2, porb2, size_t(nk2 * n), "copy POrb(2)"))) return rc;
  if ((rc = coulomb(c, nden, dD, np, dJ, np, nullptr))) return rc;
  const int32_t nocc[2] = {nk1, nk2};
  const double *cptr[2] = {dC1, dC2};
  if (tau) {  // the reference's LK screening (CASSCF list cut), FP64
    if (precision != MJK_PREC_FP64)
      return set_error(c, MJK_ERR_ARG, "cho_lk_casscf: LK screening computes in FP64 only");
    if ((rc = lk_exchange(c, nden, nocc, cptr, tau, thrv, true, dK, np))) return rc;
  } else if ((rc = exchange(c, nden, nocc, cptr, dK, np, precision))) {
    return rc;
  }
  const int32_t jsrc[2] = {0, 1};
  const double fj[2] = {1.0, 1.0};        // FactCI, FactCA (cho_fmcscf.F90:71)
  const double fk[2] = {fk_i, fk_a};      // FactXI, FactXA
  if ((rc = fock_update(c, nden, dF, np, dJ, np, jsrc, fj, dK, np, fk))) return rc;
  if (do_puvx) {
    if ((rc = h2d(dC3, porb3, size_t(naorb * n), "copy POrb(3)")) ||
        (rc = h2d(dCMO, cmo, size_t(n * n), "copy CMO")))
      return rc;
    if ((rc = mjk_puvx(c, naorb, dC3, dCMO, nfro, norb, dW))) return rc;
    if ((rc = hip_check(c, hipMemcpyAsync(w_pwxy, dW, sizeof(double) * nw,
                                          hipMemcpyDeviceToHost, c->stream), "copy W_PWXY")))
      return rc;
  }
  for (int d = 0; d < nden; ++d)
    if ((rc = hip_check(c, hipMemcpyAsync(flt[d], dF + d * np, sizeof(double) * np,
                                          hipMemcpyDeviceToHost, c->stream), "copy FLT back")))
      return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "cho_fmcscf sync");
}

namespace {
// CASSCF drop-ins with point-group symmetry on the C1 engine: FI/FA and DI/DA
// irrep blocks, the orbital sets, the active MOs and CMO are embedded in the
// full basis, the C1 build runs, and FI/FA and (pw|xy) are read back into the
// irrep layout.  Orbital counts are concatenated irrep by irrep: irrep s's k-th
// orbital of a set with per-irrep offsets ok is the set's orbital ok[s] + k.
struct CasSym {
  SymMap m;
  int64_t oa[9] = {0}, oo[9] = {0};  // active-orbital and MO offsets per irrep
  int64_t NA = 0, NO = 0;
};

// A2(nk(s), nBas(s)) blocks back to back (orbital index fastest) -> the
// orbital-major Nk x n matrix of the full basis
std::vector<double> embed_rows(const SymMap &m, const double *src, const int64_t *ok) {
  const int64_t Nk = ok[m.nsym], n = m.off[m.nsym];
  std::vector<double> dst(size_t(std::max<int64_t>(Nk, 1) * n), 0.0);
  if (!src) return dst;
  const double *ps = src;
  for (int s = 0; s < m.nsym && Nk > 0; ++s) {
    const int64_t nk = ok[s + 1] - ok[s], nb = m.off[s + 1] - m.off[s];
    for (int64_t b = 0; b < nb; ++b)
      for (int64_t k = 0; k < nk; ++k) dst[size_t((ok[s] + k) + Nk * (m.off[s] + b))] = ps[k + nk * b];
    ps += nk * nb;
  }
  return dst;
}

// square A2(nBas(s), nBas(s)) blocks back to back (AO index fastest), their
// first ok[s+1] - ok[s] columns -> columns ok[s].. of the full n x n square
std::vector<double> embed_cols(const SymMap &m, const double *src, const int64_t *ok) {
  const int64_t n = m.off[m.nsym];
  std::vector<double> dst(size_t(n * n), 0.0);
  if (!src) return dst;
  const double *ps = src;
  for (int s = 0; s < m.nsym; ++s) {
    const int64_t nk = ok[s + 1] - ok[s], nb = m.off[s + 1] - m.off[s];
    for (int64_t k = 0; k < nk; ++k)
      for (int64_t a = 0; a < nb; ++a) dst[size_t((m.off[s] + a) + n * (ok[s] + k))] = ps[a + nb * k];
    ps += nb * nb;
  }
  return dst;
}

int cas_sym_setup(mjk_context *c, int32_t nsym, const int32_t *nbas, const int32_t *naorb,
                  const int32_t *nfro, const int32_t *norb, CasSym *cs, const char *who) {
  int rc;
  if ((rc = sym_map(c, nsym, nbas, &cs->m, who))) return rc;
  if (!naorb || !nfro || !norb) return set_error(c, MJK_ERR_ARG, std::string(who) + ": null pointer");
  for (int s = 0; s < nsym; ++s) {
    if (naorb[s] < 0 || naorb[s] > nbas[s] || nfro[s] < 0 || norb[s] < 0 || nfro[s] + norb[s] > nbas[s])
      return set_error(c, MJK_ERR_ARG, std::string(who) + ": bad orbital count");
    cs->oa[s + 1] = cs->oa[s] + naorb[s];
    cs->oo[s + 1] = cs->oo[s] + norb[s];
  }
  cs->NA = cs->oa[nsym];
  cs->NO = cs->oo[nsym];
  return MJK_OK;
}

// CMO%SB(s)%A2(nBas(s), nBas(s)) back to back: columns nFro(s) ..
// nFro(s) + nOrb(s) - 1 of irrep s become columns oo(s).. of the C1 CMO(n, n)
std::vector<double> embed_cmo(const CasSym &cs, const double *cmo, const int32_t *nfro) {
  const SymMap &m = cs.m;
  const int64_t n = m.off[m.nsym];
  std::vector<double> cf(size_t(n * n), 0.0);
  if (!cmo) return cf;
  const double *ps = cmo;
  for (int s = 0; s < m.nsym; ++s) {
    const int64_t nb = m.off[s + 1] - m.off[s];
    for (int64_t p = 0; p < cs.oo[s + 1] - cs.oo[s]; ++p)
      for (int64_t a = 0; a < nb; ++a) cf[size_t((m.off[s] + a) + n * (cs.oo[s] + p))] = ps[a + nb * (nfro[s] + p)];
    ps += nb * nb;
  }
  return cf;
}

// (pw|xy) of the embedded active space, wf[npw * (Y(Y+1)/2 + X) + P + NO W],
// into the symmetry-blocked order of CHO_eval_waxy
// (src/fock_util/cho_eval_waxy.F90:86-141): blocks (sp, sw, sx <= sy) with
// sp x sw = sx x sy, stacked in the loop order sp, sw, sy; inside a block p
// fastest, then w, then the xy pair: y(y+1)/2 + x on a diagonal block, x +
// nA(sx) y otherwise.  Irreps are 0-based here, so the D2h product is XOR.
void waxy_blocks(const CasSym &cs, const int32_t *naorb, const int32_t *norb, const double *wf,
                 double *w_pwxy) {
  const int nsym = cs.m.nsym;
  const int64_t npw = cs.NO * cs.NA;
  int64_t stack = 0;
  for (int sp = 0; sp < nsym; ++sp)
    for (int sw = 0; sw < nsym; ++sw)
      for (int sy = 0; sy < nsym; ++sy) {
        const int sx = sp ^ sw ^ sy;
        if (sx > sy) continue;
        const int64_t nxs = naorb[sx], nys = naorb[sy], np_ = norb[sp], nw = naorb[sw];
        const int64_t pairs = sx == sy ? nys * (nys + 1) / 2 : nxs * nys;
        for (int64_t y = 0; y < nys; ++y)
          for (int64_t x = 0; x < (sx == sy ? y + 1 : nxs); ++x) {
            const int64_t ixy = sx == sy ? y * (y + 1) / 2 + x : x + nxs * y;
            const int64_t X = cs.oa[sx] + x, Y = cs.oa[sy] + y;  // X <= Y
            const double *src = wf + npw * (Y * (Y + 1) / 2 + X);
            double *dst = w_pwxy + stack + np_ * nw * ixy;
            for (int64_t w = 0; w < nw; ++w)
              for (int64_t p = 0; p < np_; ++p) dst[p + np_ * w] = src[(cs.oo[sp] + p) + cs.NO * (cs.oa[sw] + w)];
          }
        stack += np_ * nw * pairs;
      }
}
}  // namespace

extern "C" {

int mjk_cho_fmcscf(mjk_context *c, double *const *flt, int32_t nforb, int32_t niorb, int32_t naorb,
                   double factxi, const double *const *dlt, int32_t do_active, const double *porb1,
                   int32_t nchm, const double *porb2, const double *porb3, double *w_pwxy,
                   const double *cmo, int32_t nfro, int32_t norb, double exfac,
                   int32_t precision) {
  (void)exfac;  // the reference only prints a warning for ExFac /= 1 (cho_fmcscf.F90:88-94)
  if (!c) return MJK_ERR_ARG;
  return fmcscf_core(c, flt, nforb, niorb, naorb, factxi, -0.5, dlt, do_active, porb1, nchm, porb2,
                     porb3, w_pwxy, cmo, nfro, norb, precision);  // FactXA = -1/2, cho_fmcscf.F90:74
}

int mjk_cho_fmcscf_sym(mjk_context *c, int32_t nsym, const int32_t *nbas, double *const *flt,
                       const int32_t *nforb, const int32_t *niorb, const int32_t *naorb,
                       double factxi, const double *const *dlt, int32_t do_active,
                       const double *porb1, const int32_t *nchm, const double *porb2,
                       const double *porb3, double *w_pwxy, const double *cmo,
                       const int32_t *nfro, const int32_t *norb, double exfac, int32_t precision) {
  if (!c) return MJK_ERR_ARG;
  if (!flt || !dlt || !nforb || !niorb || !nchm)
    return set_error(c, MJK_ERR_ARG, "cho_fmcscf_sym: null pointer");
  CasSym cs;
  int rc;
  if ((rc = cas_sym_setup(c, nsym, nbas, naorb, nfro, norb, &cs, "cho_fmcscf_sym"))) return rc;
  const SymMap &m = cs.m;
  const int64_t np = c->np;
  int64_t oi[9] = {0}, oc[9] = {0};
  for (int s = 0; s < nsym; ++s) {
    if (nforb[s] < 0 || niorb[s] < 0 || nforb[s] + niorb[s] > nbas[s] || nchm[s] < 0 || nchm[s] > nbas[s])
      return set_error(c, MJK_ERR_ARG, "cho_fmcscf_sym: bad orbital count");
    oi[s + 1] = oi[s] + nforb[s] + niorb[s];
    oc[s + 1] = oc[s] + (do_active ? nchm[s] : 0);
  }
  const int nd = do_active ? 2 : 1;
  for (int d = 0; d < nd; ++d)
    if (!flt[d] || !dlt[d]) return set_error(c, MJK_ERR_ARG, "cho_fmcscf_sym: null FLT(d) or DLT(d)");
  if ((oi[nsym] > 0 && !porb1) || (oc[nsym] > 0 && !porb2) || (cs.NA > 0 && (!porb3 || !w_pwxy || !cmo)))
    return set_error(c, MJK_ERR_ARG, "cho_fmcscf_sym: null orbital array");
  std::vector<double> ff(size_t(2 * np), 0.0), df(size_t(2 * np), 0.0);
  m.each([&](int64_t t, int64_t g) {
    for (int d = 0; d < nd; ++d) {
      ff[size_t(d * np + g)] = flt[d][t];
      df[size_t(d * np + g)] = dlt[d][t];
    }
  });
  const std::vector<double> p1 = embed_rows(m, porb1, oi), p2 = embed_rows(m, porb2, oc),
                            p3 = embed_rows(m, porb3, cs.oa), cf = embed_cmo(cs, cmo, nfro);
  std::vector<double> wf(size_t(std::max<int64_t>(cs.NO * cs.NA * (cs.NA * (cs.NA + 1) / 2), 1)), 0.0);
  double *fp[2] = {ff.data(), ff.data() + np};
  const double *dp[2] = {df.data(), df.data() + np};
  if ((rc = mjk_cho_fmcscf(c, fp, 0, int32_t(oi[nsym]), int32_t(cs.NA), factxi, dp, do_active, p1.data(),
                           int32_t(oc[nsym]), p2.data(), p3.data(), cs.NA > 0 ? wf.data() : nullptr,
                           cf.data(), 0, int32_t(cs.NO), exfac, precision)))
    return rc;
  m.each([&](int64_t t, int64_t g) {
    for (int d = 0; d < nd; ++d) flt[d][t] = ff[size_t(d * np + g)];
  });
  if (cs.NA > 0) waxy_blocks(cs, naorb, norb, wf.data(), w_pwxy);
  return MJK_OK;
}

int mjk_cho_lk_casscf(mjk_context *c, const double *const *dlt, double *const *flt, const double *msq,
                      double *w_pwxy, double factxi, int32_t nfiorb, int32_t naorb, int32_t nchm,
                      const double *ash1, const double *ash2, int32_t do_active, int32_t nscreen,
                      double dmpk, double dfmat, const double *cmo, int32_t nfro, int32_t norb,
                      double exfac, int32_t precision) {
  // Without mjk_set_lk_screening every shell pair and orbital is computed (the
  // reference's result for dmpk -> 0); with it, the reference's LK screening.
  if (!c) return MJK_ERR_ARG;
  if (dmpk < 0.0) return set_error(c, MJK_ERR_ARG, "cho_lk_casscf: dmpk < 0");
  if (c->lk_on) {
    const int rc = lk_ready(c, 0, nullptr, "cho_lk_casscf");
    if (rc) return rc;
  }
  const int64_t n = c->n;
  if (nfiorb < 0 || nfiorb > n || nchm < 0 || nchm > n)
    return set_error(c, MJK_ERR_ARG, "cho_lk_casscf: bad orbital count");
  const int32_t nk2 = do_active ? nchm : 0;
  if ((nfiorb > 0 && !msq) || (nk2 > 0 && !ash2))
    return set_error(c, MJK_ERR_ARG, "cho_lk_casscf: null orbital matrix");
  // MSQ and Ash(2) are square A2(nBas,nBas) with the AO index fastest; the
  // engine reads the orbital-major A2(nk,nBas) layout
  std::vector<double> p1(static_cast<size_t>(nfiorb) * n), p2(static_cast<size_t>(nk2) * n);
  for (int64_t k = 0; k < nfiorb; ++k)
    for (int64_t b = 0; b < n; ++b) p1[size_t(k + nfiorb * b)] = msq[b + n * k];
  for (int64_t k = 0; k < nk2; ++k)
    for (int64_t b = 0; b < n; ++b) p2[size_t(k + nk2 * b)] = ash2[b + n * k];
  double tau[2] = {0.0, 0.0}, thrv[2] = {0.0, 0.0};
  if (c->lk_on) {
    const int32_t no[2] = {nfiorb, nk2};
    lk_thresholds(c, do_active ? 2 : 1, no, dmpk, dfmat, tau, thrv);
    c->lk_nscreen = nscreen;
  }
  // FactX = [FactXI*ExFac, -1/2*ExFac], cho_lk_casscf.F90:113
  return fmcscf_core(c, flt, 0, nfiorb, naorb, factxi * exfac, -0.5 * exfac, dlt, do_active,
                     p1.data(), nchm, p2.data(), ash1, w_pwxy, cmo, nfro, norb, precision,
                     c->lk_on ? tau : nullptr, c->lk_on ? thrv : nullptr);
}

int mjk_cho_lk_casscf_sym(mjk_context *c, int32_t nsym, const int32_t *nbas, const double *const *dlt,
                          double *const *flt, const double *msq, double *w_pwxy, double factxi,
                          const int32_t *nchi, const int32_t *naorb, const int32_t *nchm,
                          const double *ash1, const double *ash2, int32_t do_active, int32_t nscreen,
                          double dmpk, double dfmat, const double *cmo, const int32_t *nfro,
                          const int32_t *norb, double exfac, int32_t precision) {
  if (!c) return MJK_ERR_ARG;
  if (!flt || !dlt || !nchi || !nchm) return set_error(c, MJK_ERR_ARG, "cho_lk_casscf_sym: null pointer");
  if (c->lk_on && nsym > 1)  // the reference screens per irrep shell blocks
    return set_error(c, MJK_ERR_ARG, "cho_lk_casscf_sym: LK screening with nSym > 1 is not supported");
  CasSym cs;
  int rc;
  if ((rc = cas_sym_setup(c, nsym, nbas, naorb, nfro, norb, &cs, "cho_lk_casscf_sym"))) return rc;
  const SymMap &m = cs.m;
  const int64_t np = c->np;
  int64_t oi[9] = {0}, oc[9] = {0};
  for (int s = 0; s < nsym; ++s) {
    if (nchi[s] < 0 || nchi[s] > nbas[s] || nchm[s] < 0 || nchm[s] > nbas[s])
      return set_error(c, MJK_ERR_ARG, "cho_lk_casscf_sym: bad orbital count");
    oi[s + 1] = oi[s] + nchi[s];
    oc[s + 1] = oc[s] + (do_active ? nchm[s] : 0);
  }
  const int nd = do_active ? 2 : 1;
  for (int d = 0; d < nd; ++d)
    if (!flt[d] || !dlt[d]) return set_error(c, MJK_ERR_ARG, "cho_lk_casscf_sym: null FLT(d) or DLT(d)");
  if ((oi[nsym] > 0 && !msq) || (oc[nsym] > 0 && !ash2) || (cs.NA > 0 && (!ash1 || !w_pwxy || !cmo)))
    return set_error(c, MJK_ERR_ARG, "cho_lk_casscf_sym: null orbital array");
  std::vector<double> ff(size_t(2 * np), 0.0), df(size_t(2 * np), 0.0);
  m.each([&](int64_t t, int64_t g) {
    for (int d = 0; d < nd; ++d) {
      ff[size_t(d * np + g)] = flt[d][t];
      df[size_t(d * np + g)] = dlt[d][t];
    }
  });
  // MSQ%SB(s) and CVa(2)%SB(s) are square A2(nBas(s),nBas(s)), AO index
  // fastest; CVa(1)%SB(s) is A2(nAorb(s),nBas(s))
  const std::vector<double> sq1 = embed_cols(m, msq, oi), sq2 = embed_cols(m, do_active ? ash2 : nullptr, oc),
                            a1 = embed_rows(m, ash1, cs.oa), cf = embed_cmo(cs, cmo, nfro);
  std::vector<double> wf(size_t(std::max<int64_t>(cs.NO * cs.NA * (cs.NA * (cs.NA + 1) / 2), 1)), 0.0);
  double *fp[2] = {ff.data(), ff.data() + np};
  const double *dp[2] = {df.data(), df.data() + np};
  if ((rc = mjk_cho_lk_casscf(c, dp, fp, sq1.data(), cs.NA > 0 ? wf.data() : nullptr, factxi,
                              int32_t(oi[nsym]), int32_t(cs.NA), int32_t(oc[nsym]), a1.data(), sq2.data(),
                              do_active, nscreen, dmpk, dfmat, cf.data(), 0, int32_t(cs.NO), exfac,
                              precision)))
    return rc;
  m.each([&](int64_t t, int64_t g) {
    for (int d = 0; d < nd; ++d) flt[d][t] = ff[size_t(d * np + g)];
  });
  if (cs.NA > 0) waxy_blocks(cs, naorb, norb, wf.data(), w_pwxy);
  return MJK_OK;
}

}  // extern "C"
