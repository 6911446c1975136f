// Internal declarations of the MI355X Cholesky J/K library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cstdint>
#include <string>
#include <vector>

#include "molcas_jk.h"

namespace mjk {

typedef double v2d __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int64_t kRowAlign = 512;   // L row stride granularity (doubles, 4 KiB)
constexpr int kTileA = 64;           // AO tile of the exchange kernels
constexpr int kNumPhases = MJK_NUM_PHASES;

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
inline int64_t ceil_div(int64_t x, int64_t m) { return (x + m - 1) / m; }

// Grow-only device buffer owned by a context.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// Device workspaces.  Contexts created with mjk_create_shared use the same
// Workspace and must queue their work on the same stream.
struct Workspace {
  DevBuf d, vpart, v, jpart, cbank, cb32, cb3, x, kpart, stage, j, k, red, misc, sets, xn2, cnorm, gb32;
  ~Workspace() {
    for (DevBuf *b : {&d, &vpart, &v, &jpart, &cbank, &cb32, &cb3, &x, &kpart, &stage, &j, &k, &red,
                      &misc, &sets, &xn2, &cnorm, &gb32})
      if (b->ptr) (void)hipFree(b->ptr);
  }
};

// Tile screening of the exchange kernels (all pointers null = off).
struct Screen {
  const double *lnorm = nullptr;      // [J][strip] ||L^J[:, 64-AO strip]||_F
  const double *cmax = nullptr;       // device scalar: max over sets of ||C_d||_F
  double thr_x = 0.0;                 // half-transform tile threshold
  const double *xn2 = nullptr;        // [d][J][strip] squared norms of the chunk's X blocks
  double tau = 0.0;                   // K tile threshold
  int nstrip = 0;                     // n_pad / 64
  unsigned long long *cnt = nullptr;  // [K items skipped, X tiles skipped]
};

// Coulomb pass V = L^T D fused into the FP64 half-transform (p null = off):
// workgroup (vector J, AO tile t) adds sum_{a in t} sum_{b >= a} L^J_ba D_ba
// into vpart[t * m + J], with D the folded density in packed storage, so
// sum_t vpart[t * m + J] = sum_{b >= a} L^J_ba D_ba.
struct FuseV {
  const double *p = nullptr;
  double *vpart = nullptr;
  int64_t m = 0;   // row stride of vpart (vector capacity)
  int ntile = 0;   // AO tiles per vector of the last launch (set by the launcher)
};

// Z^J = C^T L^J C saved by the exchange in progress for the backward (z null =
// off; one orbital set of nocc columns): nocc x nocc row-major per vector, at
// z + J * nocc * nocc for every vector J of the call.
struct ZSave {
  double *z = nullptr;
  int32_t nocc = 0;
};

struct ReducedSet {
  int64_t nrs = 0;
  bool identity = false;      // rs position r maps to LT index r for all r
  int32_t *d_map = nullptr;   // device copy of rs -> LT map (0-based)
  std::vector<int32_t> h_map; // host copy (shell-pair occupancy of the LK screening)
};

// State of the reference-form LK screening that the reference keeps in module
// globals (use Cholesky: nShell, nBasSh, the diagonal CHO_IODIAG reads, MaxRed;
// Fock_util_global: Estimate, Update) and the device work buffers of one call.
struct LKState {
  int32_t nshell = 0;
  std::vector<int32_t> nbassh;            // host
  int64_t version = 0;                    // bumped when the shells change
  DevBuf shell, koff, nbsh;               // device: shell of each AO, first AO, size
  bool diag_set = false;
  DevBuf diag0;                           // [np] the diagonal at the start of a call
  int32_t maxred = 0, estimate = 0, update = 1;
  std::vector<DevBuf> present;            // per reduced set: [nshp] shell pair occupied
  std::vector<int64_t> present_version;
  DevBuf present_all;                     // every shell pair (vectors written in place)
  int64_t present_all_version = -1;
  // per call
  DevBuf diag, dh, ylk, svshp, sumac, mlk, listed, keep, faa, x, kacc, stats;
  long long last_stats[3] = {0, 0, 0};    // listed shells, half-transformed shell pairs, K blocks
  int64_t last_batches = 0;
  ~LKState() {
    for (DevBuf *b : {&shell, &koff, &nbsh, &diag0, &present_all, &diag, &dh, &ylk, &svshp, &sumac,
                      &mlk, &listed, &keep, &faa, &x, &kacc, &stats})
      if (b->ptr) (void)hipFree(b->ptr);
    for (DevBuf &b : present)
      if (b.ptr) (void)hipFree(b.ptr);
  }
};

struct Context {
  int32_t n = 0;          // nBas
  int64_t np = 0;         // n(n+1)/2
  int64_t ldl = 0;        // L row stride (doubles)
  int32_t n_pad = 0;      // n rounded up to kTileA
  int64_t nchol = 0;      // vector capacity
  double *L = nullptr;    // nchol x ldl, packed LT per vector, zero padded
  // per vector: 0 = not uploaded, r + 1 = uploaded on reduced set r, -1 = written in
  // place on the full LT (mjk_device_vectors)
  std::vector<int32_t> vset;
  int64_t nloaded = 0;
  std::vector<ReducedSet> rsets;

  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  int32_t ncu = 256;      // compute units of the device (SYRK work-item rounds)

  int64_t x_budget = int64_t(16) << 30;  // bytes for one chunk of X

  // precision ladder of MJK_PREC_LADDER (mjk_set_ladder); -1 = unset (FP64)
  int64_t lad_fp32 = -1, lad_bf16 = -1;

  // workspaces
  DevBuf ws_l32;          // FP32 shadow of L (mixed-precision exchange)
  bool l32_valid = false;
  DevBuf lnorm;           // ||L^J[:, 64-AO strip]||_F per vector and strip (screening)
  bool lnorm_valid = false;

  // LK screening of the LK drop-ins (mjk_set_lk_screening) and its counters
  bool lk_on = false;
  double lk_thrcom = 1.0e-6;
  int32_t lk_nscreen = 10;        // nScreen of the LK call in progress
  LKState lk;
  DevBuf scount;                  // device: [K items skipped, X tiles skipped]
  Screen scr;                     // screening of the exchange launch in progress
  FuseV fv;                       // Coulomb V fused into the exchange in progress
  ZSave zs;                       // Z^J = C^T L^J C kept by the exchange in progress
  int64_t k_items = 0, x_tiles = 0;  // work items of the last screened exchange
  std::shared_ptr<Workspace> wsp;  // workspaces, possibly shared with other contexts
  Workspace *w = nullptr;

  // tunables (0 = automatic); set by mjk_set_tuning or the autotuner
  int32_t tile_syrk = 0;     // SYRK tile: 64 or 128
  int32_t split_target = 0;  // SYRK workgroups per launch to aim for
  int32_t ht_ta = 0;         // half-transform AO tile: 64 or 128

  // profiling: per-phase accumulated milliseconds and launch counts
  bool profile = false;
  double phase_ms[kNumPhases] = {};
  int64_t phase_calls[kNumPhases] = {};
  std::vector<hipEvent_t> ev_pool;   // reusable events
  size_t ev_used = 0;
  int ev_open = -1;                   // pool index of the pending begin event
  struct Pending { int phase; int ev_begin; int ev_end; };
  std::vector<Pending> pending;       // recorded, not yet resolved

  std::string err;
};

// Returns a device buffer of at least `bytes`, reallocating (grow-only).
int ensure(Context *c, DevBuf &b, size_t bytes);

// Phase timing helpers (no-ops unless c->profile).  Events are recorded on the
// context stream and resolved lazily, so timing does not serialise the host.
void phase_begin(Context *c);
void phase_end(Context *c, int phase);
int resolve_phases(Context *c);

int set_error(Context *c, int rc, const std::string &msg);

// Zero-fills device memory with a kernel (no hipMemsetAsync, so the call can
// sit inside a captured HIP graph).
int fill_zero(Context *c, void *ptr, size_t bytes);
int hip_check(Context *c, hipError_t e, const char *what);

// ---- kernels (launchers) -------------------------------------------------
// Coulomb: jlt(d, p) = sum_J L(J,p) * sum_q L(J,q) dlt(d,q), for nd densities.
int coulomb(Context *c, int32_t nd, const double *dlt, int64_t ldd, double *jlt,
            int64_t ldj, double *v_out /* optional [M][nd] */);

// coulomb() with every scratch buffer inside one fixed 64 MiB device array
// (allocation-free: no hipMalloc / hipFree on any path); MJK_ERR_ARG when the
// build needs more.  Calls that overlap on one GPU share the array and race.
int coulomb_noalloc(Context *c, int32_t nd, const double *dlt, int64_t ldd, double *jlt,
                    int64_t ldj, double *v_out);

// Coulomb (one density) and FP64 exchange in one pass over the vectors: the
// V = L^T D pass runs inside the half-transform, which reads every L^J anyway;
// then the J = L V pass.  Same results as coulomb() followed by exchange().
int coulomb_exchange(Context *c, const double *dlt, int64_t ldd, double *jlt, int64_t ldj,
                     int32_t nk, const int32_t *nocc, const double *const *corb, double *klt,
                     int64_t ldk, double *v_out /* optional [M] */);

// Exchange: klt(d) = sum_J sum_k X_d(J,k,a) X_d(J,k,b) (LT), X_d = C_d^T L^J.
// corb[d] is (nocc[d] x n) with the orbital index fastest, element (k,a) at
// corb[d][k + a*nocc[d]].
int exchange(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
             double *klt, int64_t ldk, int32_t precision);

// Orbital bank: concatenates the nk sets into ws_cbank as Cb[a][k] with ldc =
// 16*ot columns; set d starts at column off[d] (multiple of 4), opad[d] =
// round_up(nocc[d], 4).  tot = sum opad (0: nothing packed).
int pack_bank(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
              std::vector<int> &off, std::vector<int> &opad, int *ot, int *tot);
// Half-transform of vectors j0..j0+jn-1 against the bank (16*ot columns):
// X[(jr*16*ot + k)*n_pad + a] = sum_b Cb[b][k] L^{j0+jr}_{ab}.
void ht_launch(Context *c, int ot, const double *cb, int64_t j0, int64_t jn, double *X);

// Z[J] = X[J] C for the jn vectors of one half-transform chunk: X holds
// 16*ot = ldc rows of n_pad per vector, FP64 (x_f32 false) or FP32; C is the
// FP64 bank (cbank).  Writes z + J * nocc * nocc, nocc x nocc row-major.
int z_from_x(Context *c, const void *X, bool x_f32, int ldc, int nocc, int64_t jn, double *z);

// Exchange with tile screening (thresholds 0 = off): a K tile of a vector
// slice is skipped when sum_J ||X^J_A||_F ||X^J_B||_F < tau, a half-transform
// tile when ||L^J[:, tile]||_F max_d ||C_d||_F < thr_x.
int exchange_screened(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                      double *klt, int64_t ldk, int32_t precision, double tau, double thr_x);

// w[J] = squared Frobenius norm of the symmetric matrix L^J (device, M doubles).
int vector_norms(Context *c, double *w);

// Exchange with vectors [0, j_fp32) in FP64, [j_fp32, j_bf16) in FP32 and
// [j_bf16, nchol) in split-BF16; all partial sums reduced in FP64.
int exchange_mixed(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                   double *klt, int64_t ldk, int64_t j_fp32, int64_t j_bf16);

// Reference-form LK exchange (CHO_LK_SCF / CHO_LK_CASSCF screening) of nd <= 2
// densities: klt(d) (plain LT, row stride ldk) = the screened sum_{J,k} X X^T,
// unscaled.  corb[d] is orbital-major (nocc[d] x n) on the device; tau/thrv per
// density; cas_cut selects the CASSCF shell-list cut MLk >= sqrt(tau).
int lk_exchange(Context *c, int32_t nd, const int32_t *nocc, const double *const *corb,
                const double *tau, const double *thrv, bool cas_cut, double *klt, int64_t ldk);

// Fock assembly: flt(d) += fj[d]*jlt(jsrc[d]) + fk[d]*klt(d).
int fock_update(Context *c, int32_t nd, double *flt, int64_t ldf, const double *jlt,
                int64_t ldj, const int32_t *jsrc, const double *fj, const double *klt,
                int64_t ldk, const double *fk);

// Energy: e = sum_p h(p) d(p) + 1/2 sum_p g(p) d(p) over LT storage.
int energy(Context *c, const double *h, const double *dlt, const double *g,
           double *e_out_dev);

}  // namespace mjk

struct mjk_context : public mjk::Context {};
