/*
 * molcas_jk.h -- C ABI of the MI355X (gfx950) Cholesky J/K Fock build.
 *
 * Drop-in for OpenMolcas's Cholesky-vector Coulomb/exchange build
 * (CHO_FSCF / CHO_FMCSCF in src/fock_util, fed by the reduced-set vectors of
 * src/cholesky_util).  All arguments are plain pointers and sizes; every entry
 * point returns an int rc (0 = success) and never aborts.  On a nonzero rc,
 * mjk_last_error() describes the failure.
 *
 * Storage conventions (nSym = 1, JSYM = 1, the C1 case):
 *   - LT packed: element (a,b), a >= b, 0-based, at a*(a+1)/2 + b
 *     (src/misc_util/index_functions.F90:138-146).
 *   - Coulomb densities are folded (off-diagonal doubled,
 *     src/misc_util/fold.F90:14); Fock matrices are plain LT.
 *   - Orbital matrices follow Porb(d)%SB(1)%A2(nOcc,nBas)
 *     (src/system_util/data_structures.F90:62-70): element (k,a) at k + a*nOcc.
 *
 * Threading: one context per thread; calls on one context are serialised by
 * the caller, as the reference's single-threaded callers do.
 */
#ifndef MOLCAS_JK_H
#define MOLCAS_JK_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mjk_context mjk_context;

/* Return codes.  33 and 77 keep the reference's meanings
 * (src/fock_util/cho_fscf.F90:147-155 and :193-196). */
enum {
  MJK_OK = 0,
  MJK_ERR_ARG = 1,        /* invalid argument / shape */
  MJK_ERR_HIP = 2,        /* HIP runtime failure */
  MJK_ERR_STATE = 3,      /* call out of order (e.g. no reduced set) */
  MJK_ERR_IO = 4,         /* vector file could not be opened or read */
  MJK_ERR_NOMEM = 33,     /* device memory exhausted */
  MJK_ERR_SHORT_READ = 77 /* vectors requested that were never uploaded */
};

/* Arithmetic of the exchange (K) contraction.  Coulomb, reductions over J and
 * the Fock/energy assembly are always FP64. */
enum {
  MJK_PREC_FP64 = 0,   /* FP64 MFMA */
  MJK_PREC_FP32 = 1,   /* FP32 MFMA, FP32 accumulate, FP64 reduction over J */
  MJK_PREC_BF16X3 = 2, /* split-BF16, 3 products, FP32 accumulate */
  MJK_PREC_LADDER = 3  /* FP64 / FP32 / split-BF16 by vector index, split
                          points of mjk_set_ladder (FP64 until set) */
};

/* Phases reported by mjk_phase_times(). */
enum {
  MJK_PHASE_COULOMB_V = 0, /* V_J = sum_pq L^J_pq D_pq           */
  MJK_PHASE_COULOMB_J = 1, /* J_pq = sum_J L^J_pq V_J            */
  MJK_PHASE_HALF_TRANSFORM = 2, /* X^J = C^T L^J (MFMA)          */
  MJK_PHASE_EXCHANGE = 3,  /* K = sum_J X^J^T X^J (MFMA)         */
  MJK_PHASE_REDUCE = 4,    /* partial-sum reductions            */
  MJK_PHASE_ASSEMBLY = 5,  /* Fock update / energy               */
  MJK_PHASE_PACK = 6,      /* density / orbital staging          */
  MJK_PHASE_BACKWARD = 7,  /* dE/dL: W = C Z and the packed W C^T (and the Z = X C of a recompute) */
  MJK_PHASE_ZSAVE = 8,     /* Z = X C kept by a forward exchange for the backward */
  MJK_NUM_PHASES = 9
};

/* ---- lifecycle --------------------------------------------------------- */

/* Creates a context on the current HIP device for nbas basis functions and
 * room for nchol Cholesky vectors kept resident in HBM.  Replaces the vector
 * buffer set up by Cho_X_Init (src/cholesky_util/cho_x_init.F90:55-469). */
int mjk_create(mjk_context **ctx, int32_t nbas, int64_t nchol);

/* Frees every device allocation of the context.  Cho_X_Final counterpart
 * (called at src/scf/wfctl_scf.F90:1202). */
int mjk_destroy(mjk_context *ctx);

/* As mjk_create, but the new context shares the device workspaces and the
 * stream of `share` (many small systems, one set of scratch buffers; calls on
 * contexts that share must not overlap).  Library-internal design: no
 * reference counterpart. */
int mjk_create_shared(mjk_context **ctx, int32_t nbas, int64_t nchol, mjk_context *share);

/* Device pointer and row stride (doubles) of the resident vector store, for
 * callers that write the vectors in place (zero copy): row J holds the packed
 * LT of vector J in its first nbas(nbas+1)/2 entries; the rest of the row
 * must stay zero.  Marks every vector as present.  Counterpart of the vector
 * buffer that Cho_X_Init fills (src/cholesky_util/cho_x_init.F90:16-52). */
int mjk_device_vectors(mjk_context *ctx, double **L, int64_t *ldl);

/* Exchange tiling: SYRK tile (0 automatic, 64, 128), SYRK workgroups per
 * launch at most (0 automatic; whole rounds of two per CU avoid a tail round),
 * half-transform AO tile (0, 64, 128). */
int mjk_set_tuning(mjk_context *ctx, int32_t tile_syrk, int32_t split_target, int32_t ht_ta);

/* Times every exchange tiling for these orbital sets and keeps the fastest
 * in the context (best time in *best_ms); klt receives the last K. */
int mjk_autotune(mjk_context *ctx, int32_t nk, const int32_t *nocc, const double *const *corb,
                 double *klt, int64_t ldk, int32_t precision, double *best_ms);

/* Text of the last error on this context ("" if none). */
const char *mjk_last_error(const mjk_context *ctx);

/* Queues all later device work of the context on `hip_stream` (a
 * hipStream_t; NULL = the default stream).  A new context uses a stream of
 * its own.  Library-internal design: no reference counterpart. */
int mjk_set_stream(mjk_context *ctx, void *hip_stream);

/* Split points of the MJK_PREC_LADDER exchange: vectors [0, j_fp32) in
 * FP64, [j_fp32, j_bf16) in FP32, [j_bf16, nchol) in split-BF16, as
 * mjk_exchange_mixed.  Until set, MJK_PREC_LADDER computes FP64.  No
 * counterpart in the reference (its exchange is FP64 throughout,
 * src/fock_util/cho_fscf.F90:250,276); it lets the reference-shaped drop-ins
 * run the ladder. */
int mjk_set_ladder(mjk_context *ctx, int64_t j_fp32, int64_t j_bf16);

/* Bytes of device memory one chunk of half-transformed vectors may use. */
int mjk_set_x_budget(mjk_context *ctx, int64_t bytes);

/* ---- vector store (src/cholesky_util vector API) ----------------------- */

/* Declares reduced set `jred` (0-based) with nrs pairs; rs_to_lt[r] is the
 * 0-based LT index of reduced-set position r (host array).  Replaces the
 * index state that Cho_X_SetRed / CHO_GETRED load
 * (src/cholesky_util/cho_x_setred.F90:43-70, cho_getred.F90:12-78) and that
 * swap_full2rs/swap_rs2full consume (src/fock_util/swap_full2rs.F90:37-46). */
int mjk_set_reduced_set(mjk_context *ctx, int32_t jred, int64_t nrs,
                        const int64_t *rs_to_lt);

/* Uploads vectors jvec..jvec+jnum-1 (0-based) of reduced set jred straight
 * from the reference's vector file (CHVEC<iSym>: vector jvec+i starts at byte
 * mbl*iadr[i] and holds nrs doubles, nrs the dimension declared for jred).
 * iadr[i] is InfVec(jvec+i+1,3,iSym), the address CHO_VECRD1 hands to
 * DDAFILE (src/cholesky_util/cho_vecrd1.F90:93-147), and mbl the file's bytes
 * per address unit, MBL(Lu) of src/io_util/ddafile.F90:36: 8 for the word-
 * addressed file of CHO_ADRVEC = 1 (DANAME_MF_WA), 512 for the block-
 * addressed file of CHO_ADRVEC = 2 (DANAME_MF; src/cholesky_util/
 * cho_openvr.F90:46-55, src/io_util/fast_io.F90:81); mbl = 0 means 8, other
 * values are rejected.  split_bytes = 0 for one file; a unit split
 * by MOLCAS_MAXFILESIZE passes its piece size min(Max_File_Length,
 * MaxFileSize*10^6) and the pieces are read from path, path'1'..'9',
 * path'A'.. (src/io_util/mpdafile.F90:92-140).  MJK_ERR_IO if a file cannot
 * be opened or read, MJK_ERR_SHORT_READ if it ends before an addressed
 * vector does. */
int mjk_upload_vectors_file(mjk_context *ctx, const char *path, int32_t jred, int64_t jvec,
                            int64_t jnum, const int64_t *iadr, int32_t mbl, int64_t split_bytes);

/* Uploads vectors jvec..jvec+jnum-1 (0-based) stored on reduced set jred as
 * lrs(nrs, jnum), column-major -- the Lrs buffer CHO_VECRD fills
 * (src/cholesky_util/cho_vecrd.F90:12-64).  lrs is a host pointer unless
 * on_device != 0. */
int mjk_upload_vectors(mjk_context *ctx, int32_t jred, int64_t jvec, int64_t jnum,
                       const double *lrs, int32_t on_device);

/* ---- the reference's Cholesky files (Cho_X_Init, nSym = 1..8) -----------
 * Host-side reading of the files the decomposition writes, so a caller passes
 * file names instead of maps and addresses.  mjk_choinfo_read parses the
 * restart file CHORST (src/cholesky_util/cho_wrrstc.F90:30-92, read by
 * cho_x_rdrst.F90:33-190: shells, ThrCom, Cho_AdrVec, MaxRed, InfRed, NumCho,
 * InfVec) and the reduced-set index file CHRED (cho_putred.F90 / cho_getred.F90
 * :46-78: nnBstRSh and IndRed per reduced set, IndRSh and iSP2F of set 1) and
 * rebuilds the LT index of every reduced-set element the way CHO_RSTOF does
 * (cho_rstof.F90:25-89), from isoshl = the runfile array ISOSHL (1-based shell
 * of each SO, nBasT values, SOs numbered irrep by irrep).  With nSym > 1 the
 * index is into the C1 triangle of all nBasT SOs (what the *_sym drop-ins
 * use), each element is checked to lie in the irrep block it is filed under,
 * and every (irrep iSym, reduced set JRED) is a map of its own, numbered
 * (iSym-1)*MaxRed + JRED-1.  Files are word addressed with 8-byte integers.
 * No device work; safe without a GPU. */
typedef struct mjk_choinfo mjk_choinfo;

/* Reads CHORST and CHRED.  *info is always set (free it with
 * mjk_choinfo_free); on a nonzero rc mjk_choinfo_error tells why (4: a file
 * cannot be read, 1: inconsistent contents or nSym outside 1..8). */
int mjk_choinfo_read(mjk_choinfo **info, const char *chorst, const char *chred,
                     const int32_t *isoshl);
const char *mjk_choinfo_error(const mjk_choinfo *info);
int mjk_choinfo_free(mjk_choinfo *info);

/* nBasT, NumCho summed over the irreps, MaxRed, nShell, Cho_AdrVec (1 word /
 * 2 block addressed vector file) and ThrCom of the files (any pointer may be
 * NULL). */
int mjk_choinfo_dims(const mjk_choinfo *info, int32_t *nbas, int64_t *numcho, int32_t *maxred,
                     int32_t *nshell, int32_t *adrvec, double *thrcom);

/* nSym and, per irrep (arrays of 8, entries past nSym zero), nBas(iSym) and
 * NumCho(iSym) (any pointer may be NULL).  No reference counterpart: the
 * reference keeps them in module variables. */
int mjk_choinfo_symmetry(const mjk_choinfo *info, int32_t *nsym, int32_t *nbas_sym,
                         int64_t *numcho_sym);

/* Map jred = (iSym-1)*MaxRed + JRED-1 (jred = JRED-1 for nSym = 1): *nrs
 * elements and, if rs_to_lt is not NULL, the 0-based LT index of each (the map
 * mjk_set_reduced_set takes). */
int mjk_choinfo_reduced_set(const mjk_choinfo *info, int32_t jred, int64_t *nrs,
                            int64_t *rs_to_lt);

/* Column col (1..5) of InfVec(1:NumCho(iSym), col, iSym), irrep 1's vectors
 * first: 2 = the map of each vector, 1-based ((iSym-1)*MaxRed + JRED; JRED for
 * nSym = 1), 3 = its disk address in the CHVEC file of its irrep. */
int mjk_choinfo_infvec(const mjk_choinfo *info, int32_t col, int64_t *values);

/* Loads a context (nbas = nBas, capacity >= NumCho) from the files: declares
 * every reduced set, uploads the vectors from chvec (CHVEC1) at InfVec(J,3)
 * with the address unit of Cho_AdrVec (split_bytes as
 * mjk_upload_vectors_file), sets the shells when the AOs of each shell are
 * consecutive, and, if chodiag is not NULL, reads the diagonal file CHODIAG
 * (nnBstRT(1) doubles in reduced-set-1 order, src/cholesky_util/
 * cho_iodiag.F90:28-33) into mjk_set_lk_diag with MaxRed of the files. */
int mjk_load_choinfo(mjk_context *ctx, const mjk_choinfo *info, const char *chvec,
                     const char *chodiag, int64_t split_bytes);

/* mjk_load_choinfo for any nSym: chvec[iSym-1] is the vector file of irrep
 * iSym (CHVEC<iSym>, cho_openvr.F90), nfiles = nSym; the context has nbas =
 * nBasT and holds the vectors of every irrep, irrep 1's first, each scattered
 * into the C1 triangle by its map (the layout the *_sym drop-ins take).  No
 * reference counterpart. */
int mjk_load_choinfo_sym(mjk_context *ctx, const mjk_choinfo *info, const char *const *chvec,
                         int32_t nfiles, const char *chodiag, int64_t split_bytes);

/* ---- reference-shaped drop-ins (host arrays, synchronous) -------------- */

/* Per-density arrays.  The reference passes nDen separately allocated DSBA
 * arrays (FLT(1:nDen), Porb(1:nDen), KLT(1:nDen): src/scf/choscf_drv.F90:
 * 285-291,401-402; FI/FA and DI/DA of src/fock_util/cho_cas_drv.F90:42-43,
 * 101-102), so every per-density argument of a drop-in is an array of nden
 * pointers, element d pointing at that density's %A0 storage.  Nothing is
 * assumed about where the densities' arrays sit relative to each other. */

/* CHO_FSCF (src/fock_util/cho_fscf.F90:14), nSym = 1:
 *   FLT(d) += J[DLT(1)] - ExFac * sum_{J,k} X(k,J,a) X(k,J,b),  d < nden,
 * nden in {1,2}; flt[d] -> FLT(d)%A0 (LT, updated), porb[d] -> Porb(d)%A0
 * with nforb[d]+niorb[d] rows (A2(nOcc,nBas)); dlt -> DLT(1)%A0, the only
 * density that enters Coulomb.  Called from src/scf/choscf_drv.F90:197,420. */
int mjk_cho_fscf(mjk_context *ctx, int32_t nden, double *const *flt, const int32_t *nforb,
                 const int32_t *niorb, const double *const *porb, const double *dlt,
                 double exfac, int32_t precision);

/* CHO_FSCF (src/fock_util/cho_fscf.F90:14) with point-group symmetry,
 * nsym in 1..8 irreps of nbas[s] functions (sum = the context's nbas; the
 * basis is ordered irrep by irrep).  Storage as the reference's DSBA types
 * flatten it (src/system_util/data_structures.F90:62-70):
 *   flt, dlt   per density, the LT blocks of irreps 1..nsym back to back
 *              (FLT(d)%SB(s)%A1), sum_s nbas[s](nbas[s]+1)/2 values;
 *   nforb,niorb (8, nden) column-major, as nForb(8,nDen) / nIorb(8,nDen);
 *   porb       per density, per irrep A2(nocc(s), nbas[s]) with
 *              nocc(s) = nforb + niorb, back to back (Porb(d)%SB(s)%A2).
 * Vectors of every symmetry JSYM go into the context through reduced sets
 * whose maps index the full C1 lower triangle (global AO indices).  The
 * build runs on the C1 engine (the irrep blocks embedded in the full basis):
 * FLT(d)%SB(s) += J - ExFac K on every irrep block; K and J are block
 * diagonal because the densities and orbitals are. */
int mjk_cho_fscf_sym(mjk_context *ctx, int32_t nsym, const int32_t *nbas, int32_t nden,
                     double *const *flt, const int32_t *nforb, const int32_t *niorb,
                     const double *const *porb, const double *dlt, double exfac,
                     int32_t precision);

/* CHO_LK_SCF (src/fock_util/cho_lk_scf.F90:14) with point-group symmetry:
 * storage and embedding as mjk_cho_fscf_sym, with porb[d] the irrep blocks
 * Porb(d)%SB(s)%A2(nbas[s], nbas[s]) (AO index fastest, the first
 * nforb + niorb columns are the orbitals) back to back and plt the folded
 * total density in irrep LT blocks.  FLT(d)%SB(s) += J[PLT] - FactXI K.
 * klt is NULL or an array of nden pointers to KLT(d)%A0: per irrep s the
 * block KLT(d)%SB(s) (nbas[s](nbas[s]+1)/2 values, blocks back to back) is
 * accumulated with -FactXI K in the reference's shell-pair-blocked storage
 * of that irrep's shells (cho_lk_scf.F90:274-292), nbassh[s + nsym*sh] =
 * nBasSh(s, sh) for the nshell shells (nshell = 0: one shell per irrep, plain
 * LT order).  Runs unscreened: with LK screening switched on and nsym > 1 it
 * returns MJK_ERR_ARG (the reference screens per irrep shell block). */
int mjk_cho_lk_scf_sym(mjk_context *ctx, int32_t nsym, const int32_t *nbas, int32_t nden,
                       double *const *flt, double *const *klt, const int32_t *nforb,
                       const int32_t *niorb, const double *const *porb, const double *plt,
                       double factxi, int32_t nscreen, double dmpk, double dfmat,
                       int32_t nshell, const int32_t *nbassh, int32_t precision);

/* CHO_LK_SCF (src/fock_util/cho_lk_scf.F90:14), nSym = 1:
 *   FLT(d) += J[PLT(1)] - FactXI * sum_{J,k} L(a,J)[k] L(b,J)[k],  d < nden,
 * nden in {1,2}; flt[d] -> FLT(d)%A0; porb[d] -> Porb(d)%A0, the square
 * A2(nbas,nbas) with the AO index fastest, columns 0..nforb[d]+niorb[d]-1
 * used (the MSQ layout of src/scf/choscf_drv.F90:208); plt -> PLT(1)%A0.
 * klt is NULL or an array of nden pointers (KLT(d)%A0, nbas(nbas+1)/2
 * values each), accumulated with -FactXI*K in the reference's shell-pair-
 * blocked storage (cho_lk_scf.F90:274-292) for the nshell shells of sizes
 * nbassh (nshell = 0: one shell, i.e. plain LT order).
 * Without mjk_set_lk_screening every shell pair and orbital is kept, the
 * reference's result for dmpk -> 0.  With it, the reference's LK screening
 * runs (per-orbital shell lists from Ylk = DiaH |C|, the vector and pair
 * cuts, cho_lk_scf.F90:547-627,659-660,713-781), on the shells of
 * mjk_set_shells and the diagonal of mjk_set_lk_diag.
 * Called from src/scf/choscf_drv.F90:208,433. */
int mjk_cho_lk_scf(mjk_context *ctx, int32_t nden, double *const *flt, double *const *klt,
                   const int32_t *nforb, const int32_t *niorb, const double *const *porb,
                   const double *plt, double factxi, int32_t nscreen, double dmpk,
                   double dfmat, int32_t nshell, const int32_t *nbassh,
                   int32_t precision);

/* CHO_FOCK_DFT_RED (src/fock_util/cho_fock_dft_red.F90:14), the J-only build
 * of pure-DFT SCF: FLT = J[DLT], written, not accumulated (the first reduced
 * set's swap_rs2full runs with add = .false., which zeroes FLT,
 * swap_rs2full.F90:37-41).  Called from src/scf/choscf_drv.F90:95,302. */
int mjk_cho_fock_dft_red(mjk_context *ctx, const double *dlt, double *flt);

/* CHO_FMCSCF (src/fock_util/cho_fmcscf.F90:14), nSym = 1:
 *   FI += J[DI] + FactXI * sum X_I X_I^T      (porb1: nforb+niorb rows)
 *   FA += J[DA] - 1/2 * sum X_A X_A^T         (porb2: nchm rows, if do_active)
 *   W_PWXY(p,w,xy) = sum_J (pw|J)(xy|J)        (porb3: naorb rows)
 * flt[0] -> FLT(1)%A0 = FI, flt[1] -> FLT(2)%A0 = FA; dlt[0] -> DLT(1)%A0 =
 * DI, dlt[1] -> DLT(2)%A0 = DA (folded), the second entries read only if
 * do_active (the separately allocated FI/FA and DI/DA of
 * src/fock_util/cho_cas_drv.F90:42-43,101-102); cmo is CMO(n,n) column-major
 * and p runs over norb MOs after nfro frozen ones; w_pwxy may be NULL when
 * naorb == 0.  Called from src/fock_util/cho_cas_drv.F90:277. */
int mjk_cho_fmcscf(mjk_context *ctx, double *const *flt, int32_t nforb, int32_t niorb,
                   int32_t naorb, double factxi, const double *const *dlt, int32_t do_active,
                   const double *porb1, int32_t nchm, const double *porb2,
                   const double *porb3, double *w_pwxy, const double *cmo,
                   int32_t nfro, int32_t norb, double exfac, int32_t precision);

/* CHO_FMCSCF (src/fock_util/cho_fmcscf.F90:14) with point-group symmetry
 * (nsym irreps, basis ordered irrep by irrep, nbas[s] functions each):
 * flt[d], dlt[d] hold the irrep LT blocks back to back (FLT(d)%SB(s)%A1,
 * DLT(d)%SB(s)%A1); nforb, niorb, naorb, nchm, nfro, norb are the per-irrep
 * counts of the reference's 8-element arrays; porb1/porb2/porb3 are
 * POrb(1..3)%SB(s)%A2(nk(s), nBas(s)) back to back with nk = nforb + niorb,
 * nchm, naorb; cmo is CMO%SB(s)%A2(nBas(s), nBas(s)) back to back and p of
 * irrep s runs over its norb[s] MOs after nfro[s].  w_pwxy receives (pw|xy)
 * in the symmetry-blocked order of src/fock_util/cho_eval_waxy.F90:86-141
 * (blocks sp, sw, sx <= sy with sp x sw = sx x sy; p fastest, then w, then
 * the xy pair, LT within a diagonal block).  The vectors (every JSYM) must be
 * in the context with reduced-set maps into the C1 triangle; the build runs
 * on the C1 engine with the irrep blocks embedded.  Called from
 * src/fock_util/cho_cas_drv.F90:277 when nSym > 1. */
int mjk_cho_fmcscf_sym(mjk_context *ctx, int32_t nsym, const int32_t *nbas, double *const *flt,
                       const int32_t *nforb, const int32_t *niorb, const int32_t *naorb,
                       double factxi, const double *const *dlt, int32_t do_active,
                       const double *porb1, const int32_t *nchm, const double *porb2,
                       const double *porb3, double *w_pwxy, const double *cmo,
                       const int32_t *nfro, const int32_t *norb, double exfac,
                       int32_t precision);

/* CD_InCore (src/cholesky_util/cd_incore.F90:35-60, cd_incore_1.F90:14-105)
 * on the device: pivoted incomplete Cholesky of the n x n matrix x
 * (column-major device array, destroyed) into vec (n x mxvec, column-major
 * device array), *numcho vectors.  thr < 0 selects the default 1e-6.  *irc
 * receives the reference's code: 0 success, 101 negative diagonal below
 * -1e-8, 102 more than mxvec vectors needed, -1 mxvec <= 0.  The DECO
 * density decomposition of src/scf/choscf_drv.F90:99-135 (thr = 1e-8 * max
 * diagonal) and src/fock_util/cho_cas_drv.F90:129,227.  n <= 4096.
 * Synchronous. */
int mjk_cd_incore(mjk_context *ctx, int32_t n, double *x, double *vec, int32_t mxvec,
                  int32_t *numcho, double thr, int32_t *irc);

/* CHO_LK_CASSCF (src/fock_util/cho_lk_casscf.F90:14), nSym = 1: the LK form
 * of CHO_FMCSCF,
 *   FI += J[DI] + FactXI*ExFac * K[MSQ],  FA += J[DA] - 1/2*ExFac * K[Ash(2)]
 * (FactX of cho_lk_casscf.F90:113; FA only if do_active), W_PWXY from the
 * active MOs ash1 (orbital-major A2(naorb,nBas)) and cmo (CMO(n,n),
 * column-major, p over norb MOs after nfro).  msq and ash2 are square
 * A2(nBas,nBas) with the AO index fastest (first nfiorb / nchm columns used).
 * dlt[0], dlt[1] -> DI, DA and flt[0], flt[1] -> FI, FA (separate arrays, the
 * second entries read only if do_active).  Screening as mjk_cho_lk_scf, with
 * the CASSCF list cut MLk >= sqrt(tau) (cho_lk_casscf.F90:688).  Called from
 * src/fock_util/cho_cas_drv.F90:285. */
int mjk_cho_lk_casscf(mjk_context *ctx, const double *const *dlt, double *const *flt,
                      const double *msq, double *w_pwxy, double factxi, int32_t nfiorb,
                      int32_t naorb, int32_t nchm, const double *ash1, const double *ash2,
                      int32_t do_active, int32_t nscreen, double dmpk, double dfmat,
                      const double *cmo, int32_t nfro, int32_t norb, double exfac,
                      int32_t precision);

/* CHO_LK_CASSCF (src/fock_util/cho_lk_casscf.F90:14) with point-group
 * symmetry, unscreened (LK screening on with nsym > 1 returns MJK_ERR_ARG:
 * the reference screens per irrep shell block).  Storage as
 * mjk_cho_fmcscf_sym: flt[d], dlt[d] irrep LT blocks back to back; msq and
 * ash2 are MSQ%SB(s)%A2 and CVa(2)%SB(s)%A2(nBas(s), nBas(s)) back to back
 * (AO index fastest, the first nchi[s] / nchm[s] columns used); ash1 is
 * CVa(1)%SB(s)%A2(naorb[s], nBas(s)) back to back; cmo, nfro, norb and
 * w_pwxy as for mjk_cho_fmcscf_sym.  Called from
 * src/fock_util/cho_cas_drv.F90:285 when nSym > 1. */
int mjk_cho_lk_casscf_sym(mjk_context *ctx, int32_t nsym, const int32_t *nbas,
                          const double *const *dlt, double *const *flt, const double *msq,
                          double *w_pwxy, double factxi, const int32_t *nchi,
                          const int32_t *naorb, const int32_t *nchm, const double *ash1,
                          const double *ash2, int32_t do_active, int32_t nscreen, double dmpk,
                          double dfmat, const double *cmo, const int32_t *nfro,
                          const int32_t *norb, double exfac, int32_t precision);

/* ---- device engine (device pointers, asynchronous on the context stream) */

/* Coulomb for nd folded LT densities dlt(d*ldd + p): jlt(d*ldj + p) =
 * sum_J L^J_p V_J(d), V_J(d) = sum_q L^J_q dlt(d,q); two streaming passes over
 * the resident vectors.  v_out (optional, M x nd) receives V.
 * Replaces the DGEMV pair of src/fock_util/cho_fscf.F90:208,215. */
int mjk_coulomb(mjk_context *ctx, int32_t nd, const double *dlt, int64_t ldd,
                double *jlt, int64_t ldj, double *v_out);

/* Exchange for nk orbital sets: klt(d*ldk + LT(a,b)) = sum_{J,k}
 * X_d(k,J,a) X_d(k,J,b), X_d(k,J,a) = sum_b C_d(k,b) L^J_ab.  corb is a host
 * array of nk device pointers, corb[d] holding nocc[d] x nbas (orbital index
 * fastest).  Replaces Cho_X_getVtra + DGEMM_Tri
 * (src/fock_util/cho_fscf.F90:250,276). */
int mjk_exchange(mjk_context *ctx, int32_t nk, const int32_t *nocc,
                 const double *const *corb, double *klt, int64_t ldk,
                 int32_t precision);

/* mjk_coulomb without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it).  No counterpart in the reference.  Same kernels
 * and results as mjk_coulomb; every scratch buffer (the padded density, the
 * partial sums of V and J, and V when v_out is NULL) lives in one fixed
 * 64 MiB device array per GPU, set up with the library and held until the
 * process ends.  Builds that need more scratch than 64 MiB return
 * MJK_ERR_ARG (the scratch grows with nd, nchol and nbas(nbas+1)/2: about
 * nd * nbas(nbas+1)/2 * 8 bytes times the number of vector slices of the J
 * pass).  Calls that overlap on one GPU share the array and race; calls on
 * one context are serialised by its stream. */
int mjk_coulomb_noalloc(mjk_context *ctx, int32_t nd, const double *dlt, int64_t ldd,
                        double *jlt, int64_t ldj, double *v_out);

/* Coulomb of one folded density and FP64 exchange of nk orbital sets in one
 * pass over the vectors: jlt as mjk_coulomb (nd = 1), klt as mjk_exchange in
 * FP64.  The V_J = sum_p L^J_p D_p pass runs inside the half-transform, which
 * reads every L^J anyway, so the vectors are streamed once for V and X and
 * once for J.  v_out (optional, device, nchol values) receives V.  One vector
 * loop of CHO_FSCF: the DGEMV pair and Cho_X_getVtra + DGEMM_Tri
 * (src/fock_util/cho_fscf.F90:208,215,250,276). */
int mjk_coulomb_exchange(mjk_context *ctx, const double *dlt, int64_t ldd, double *jlt,
                         int64_t ldj, int32_t nk, const int32_t *nocc,
                         const double *const *corb, double *klt, int64_t ldk,
                         double *v_out);

/* One orbital set (1 <= nocc <= 256), and in the same pass over the vectors
 * Z^J = C^T L^J C for every vector, written to z (device, nchol * nocc * nocc
 * doubles: vector J at z + J*nocc*nocc, nocc x nocc row-major) for the
 * backward (mjk_grad_vectors_z), which then needs no second half-transform.
 * mjk_exchange_z: klt (one row) as mjk_exchange with any precision mode; Z
 * carries the precision of the segment its vector lies in (X of an FP32 or
 * split-BF16 segment is multiplied by C in FP64).  mjk_coulomb_exchange_z:
 * jlt, klt and v_out as mjk_coulomb_exchange with nk = 1.  No counterpart in
 * the reference (which has no gradient); the exchange part replaces
 * Cho_X_getVtra + DGEMM_Tri (src/fock_util/cho_fscf.F90:250,276). */
int mjk_exchange_z(mjk_context *ctx, int32_t nocc, const double *corb, double *klt,
                   int32_t precision, double *z);
int mjk_coulomb_exchange_z(mjk_context *ctx, const double *dlt, double *jlt, int32_t nocc,
                           const double *corb, double *klt, double *v_out, double *z);

/* Exchange as mjk_exchange with a precision ladder over the vector index:
 * J in [0, j_fp32) in FP64, [j_fp32, j_bf16) in FP32 and [j_bf16, nchol) in
 * split-BF16 (three products); every partial sum is reduced in FP64.  With
 * vectors ordered by decreasing norm (the Cholesky pivot order) the low
 * precisions only see the small tail.  0 <= j_fp32 <= j_bf16 <= nchol.
 * No single counterpart in the reference: it replaces the same
 * Cho_X_getVtra + DGEMM_Tri pair (src/fock_util/cho_fscf.F90:250,276). */
int mjk_exchange_mixed(mjk_context *ctx, int32_t nk, const int32_t *nocc,
                       const double *const *corb, double *klt, int64_t ldk,
                       int64_t j_fp32, int64_t j_bf16);

/* Exchange as mjk_exchange with tile screening, a tile-granular screening of
 * the device engine (the drop-ins screen the reference's way, see
 * mjk_set_lk_screening): a K tile (TM x TM AOs) of a vector slice
 * is skipped when sum_J ||X^J_A||_F ||X^J_B||_F < tau (an elementwise bound
 * of its contribution), a half-transform tile (one vector, 64 AOs) when
 * ||L^J[:, tile]||_F max_d ||C_d||_F < thr_x.  tau = thr_x = 0 computes
 * everything; precision MJK_PREC_BF16X3 cannot be screened. */
int mjk_exchange_screened(mjk_context *ctx, int32_t nk, const int32_t *nocc,
                          const double *const *corb, double *klt, int64_t ldk,
                          int32_t precision, double tau, double thr_x);

/* Work items of the last exchange call: K tiles x slices launched and
 * skipped, half-transform tiles launched and skipped (zero when the call was
 * not screened).  Synchronises the context stream.  No counterpart in the
 * reference. */
int mjk_screening_stats(mjk_context *ctx, int64_t *k_items, int64_t *k_skipped,
                        int64_t *x_tiles, int64_t *x_skipped);

/* Turns LK screening of mjk_cho_lk_scf / mjk_cho_lk_casscf on (enable != 0)
 * or off (default).  thrcom is the decomposition threshold ThrCom that
 * Cho_LK_ScreeningThreshold reads from the Cholesky module
 * (src/fock_util/cho_lk_screeningthreshold.F90:24-39).  The drop-ins then
 * screen the way the reference does, per density d:
 *   tau(d) = LKThr/max(1,nOcc_d)*dmpk (/MaxRed with Estimate),
 *   thrv(d) = sqrt(LKThr/(max(1,nOcc_d)*NumCho))*dmpk
 * (src/fock_util/cho_lk_scf.F90:143-171, cho_lk_casscf.F90:164-200); per
 * batch of 25 vectors of one reduced set (Cho_LK_MaxVecPerBatch) and per
 * orbital: shell lists from Y = sqrt|D| |C| recomputed every nScreen reduced
 * sets (:547-627), half-transform shell pairs with sqrt|SumAClk SvShp| >=
 * thrv (:659-702), K blocks with MLk MLk >= tau, sqrt|Faa Faa| >= tau/MaxRed
 * and the Keep flags (:713-830).  It needs mjk_set_shells (or the call's
 * own nbassh) and mjk_set_lk_diag; vectors must come reduced set by reduced
 * set.  Screened builds compute in FP64. */
int mjk_set_lk_screening(mjk_context *ctx, int32_t enable, double thrcom);

/* Shell structure of the basis, nBasSh(1,1:nShell) of the Cholesky module
 * (src/cholesky_util/cholesky.F90): shell s holds nbassh[s] consecutive AOs,
 * shells in order, sum = nbas.  Used by the LK screening. */
int mjk_set_shells(mjk_context *ctx, int32_t nshell, const int32_t *nbassh);

/* The diagonal the LK screening starts every call from: the integral
 * diagonal D(ab) that CHO_IODIAG reads (src/fock_util/cho_lk_scf.F90:180),
 * nbas(nbas+1)/2 values in LT order, zero for pairs outside the first
 * reduced set (host array, copied).  maxred is MaxRed (0: the number of
 * declared reduced sets); estimate / update are the Estimate / Update
 * switches of src/fock_util/fock_util_global.F90:56-59 (defaults .false. /
 * .true.): with update the diagonal loses sum_J L^2 after every batch, with
 * estimate it is rebuilt from each batch. */
int mjk_set_lk_diag(mjk_context *ctx, const double *diag, int32_t maxred, int32_t estimate,
                    int32_t update);

/* Counters of the last screened LK call: stats[0] shells listed over all
 * (screening batch, orbital), stats[1] half-transformed (batch, orbital,
 * shell pair), stats[2] K blocks computed, stats[3] batches.  No counterpart
 * in the reference (its timers count time, not work). */
int mjk_lk_stats(mjk_context *ctx, int64_t *stats);

/* w_dev[J] = ||L^J||_F^2 of the full symmetric matrix (device, nchol
 * doubles): the weights a caller uses to place the ladder split points.
 * No counterpart in the reference. */
int mjk_vector_norms(mjk_context *ctx, double *w_dev);

/* flt(d) += fj[d] * jlt(jsrc[d]) + fk[d] * klt(d), d < nd (jsrc[d] < 0: no
 * J term).  The LT update of src/fock_util/cho_fscf.F90:276,301. */
int mjk_fock_update(mjk_context *ctx, int32_t nd, double *flt, int64_t ldf,
                    const double *jlt, int64_t ldj, const int32_t *jsrc,
                    const double *fj, const double *klt, int64_t ldk,
                    const double *fk);

/* e_dev[0] = sum_LT h*dlt + 1/2 sum_LT g*dlt (FP64), the energy expression
 * of src/scf/eneclc.F90:58,77,104 without E_nuc. */
int mjk_energy(mjk_context *ctx, const double *h, const double *dlt,
               const double *g, double *e_dev);

/* (pw|xy) integrals from the resident vectors (device pointers):
 *   w(p + norb*(w + naorb*xy)) = sum_J (pw|J)(xy|J),
 * (pw|J) = sum_ab CMO(a,nfro+p) C3(w,b) L^J_ab, (xy|J) = sum_ab C3(x,a) C3(y,b)
 * L^J_ab with xy = y(y+1)/2 + x, x <= y.  porb3 is C3 (naorb x nbas, orbital
 * index fastest), cmo is CMO(nbas,nbas) column-major.  Replaces the Lxy /
 * CHO_eval_waxy steps of src/fock_util/cho_fmcscf.F90:403-487. */
int mjk_puvx(mjk_context *ctx, int32_t naorb, const double *porb3, const double *cmo,
             int32_t nfro, int32_t norb, double *w);

/* ---- energy gradient (device pointers; no reference counterpart) -------
 * For E = tr(P h) + 1/2 sum_J V_J^2 - 1/4 sum_J tr(P L^J P L^J), P = 2 C C^T
 * (the energy of src/scf/eneclc.F90 with the Fock matrix of CHO_FSCF):
 * dE/dC = 4 F C and dE/dL^J = V_J P - 1/2 P L^J P. */

/* Folded LT density dlt = fold(scale * C C^T) for C = corb (nbas x nocc,
 * orbital index fastest): the Fold step of src/misc_util/fold.F90 applied to
 * the density the SCF builds from its orbitals (src/scf/dmat.F90). */
int mjk_density(mjk_context *ctx, int32_t nocc, const double *corb, double scale,
                double *dlt);

/* dc = scale * F C, F in plain LT storage, corb and dc nbas x nocc (orbital
 * index fastest).  With scale = 4 and F = h + J - 1/2 K this is dE/dC. */
int mjk_grad_orbitals(mjk_context *ctx, int32_t nocc, const double *corb, const double *flt,
                      double scale, double *dc);

/* dl(J*ldd + LT(a,b)) = s_ab * (alpha V_J P_ab + beta (C Z^J C^T)_ab),
 * Z^J = C^T L^J C, s_ab = 2 for a != b and 1 on the diagonal (packed
 * parameters).  p_lt is P in plain LT storage, or NULL for P = 2 C C^T; v is
 * V_J (nchol values, e.g. from mjk_coulomb).  alpha = 1, beta = -2 gives
 * dE/dL.  X^J = C^T L^J is recomputed in chunks bounded by the X budget. */
int mjk_grad_vectors(mjk_context *ctx, int32_t nocc, const double *corb, const double *p_lt,
                     const double *v, double alpha, double beta, double *dl, int64_t ldd);

/* mjk_grad_vectors with Z^J = C^T L^J C given (z as written by mjk_exchange_z
 * or mjk_coulomb_exchange_z for the same orbitals): no half-transform; per
 * vector chunk W = C Z^J, then the packed dl.  No reference counterpart. */
int mjk_grad_vectors_z(mjk_context *ctx, int32_t nocc, const double *corb, const double *p_lt,
                       const double *v, const double *z, double alpha, double beta, double *dl,
                       int64_t ldd);

/* mjk_grad_vectors_z with the products C Z^J C^T of the vectors J >= j_fp32
 * formed in FP32 (FP64 V P term and output): the backward of the precision
 * ladder, whose vectors past its FP64 head carry FP32-sized errors already.
 * 0 <= j_fp32 <= nchol; j_fp32 = nchol is mjk_grad_vectors_z.  For nocc > 128,
 * or a packed dL row of 2^31 bytes or more (nbas >= 23170), every vector is
 * formed in FP64.  No reference counterpart. */
int mjk_grad_vectors_z_mixed(mjk_context *ctx, int32_t nocc, const double *corb, const double *p_lt,
                             const double *v, const double *z, double alpha, double beta, double *dl,
                             int64_t ldd, int64_t j_fp32);

/* Waits for all work queued on the context stream. */
int mjk_sync(mjk_context *ctx);

/* ---- measurement -------------------------------------------------------- */

/* Turns per-phase HIP-event timing on (1) or off (0); resets the counters. */
int mjk_set_profiling(mjk_context *ctx, int32_t on);

/* Measured peak of the current device (no reference counterpart):
 * kind 0 FP64 MFMA, 1 FP32 MFMA, 2 BF16 MFMA, 3 FP64 VALU FMA (TFLOP/s),
 * 4 device-to-device copy (TB/s, bytes read + written). */
int mjk_measure_peak(mjk_context *ctx, int32_t kind, double *value);

/* Copies accumulated milliseconds and launch counts per phase (arrays of
 * MJK_NUM_PHASES entries; either may be NULL). */
int mjk_phase_times(mjk_context *ctx, double *ms, int64_t *calls);

#ifdef __cplusplus
}
#endif
#endif /* MOLCAS_JK_H */
