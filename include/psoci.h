/*
 * psoci.h -- C ABI of the MI355X-native PSOCI Hamiltonian / Davidson-sigma path.
 *
 * One context owns one GPU's copy of a determinant/SEF space, its integrals,
 * an optional stored Hamiltonian and the Davidson work vectors.  Every entry
 * point returns an int status (PSOCI_OK or a negative PSOCI_ERR_* code) and
 * never aborts; psoci_last_error() gives the message of the last failure on
 * that context.  Orbital and configuration indices are 0-based; packed
 * integral arrays keep the reference's 1-based packing formulas (stored
 * 0-based, see psoci_set_integrals).  Host arrays passed in are copied;
 * output buffers are caller-allocated with sizes from psoci_space_info.
 *
 * Each entry point names the reference interface it replaces
 * (psociRestart-Cori/<file>:<line> in the reference tree).
 */
#ifndef PSOCI_H
#define PSOCI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psoci_ctx psoci_ctx;

enum {
  PSOCI_OK = 0,
  PSOCI_ERR_ARG = -1,         /* bad argument or size */
  PSOCI_ERR_UNSUPPORTED = -2, /* input outside the implemented model */
  PSOCI_ERR_STATE = -3,       /* call order: e.g. sigma before integrals */
  PSOCI_ERR_HIP = -4,         /* HIP runtime failure */
  PSOCI_ERR_NOMEM = -5,       /* device allocation failed */
  PSOCI_ERR_CONVERGE = -6,    /* Davidson hit its iteration limit */
  PSOCI_ERR_NUMERIC = -7      /* breakdown: linear dependence, non-SPD overlap */
};

enum { PSOCI_FP64 = 0, PSOCI_FP32 = 1, PSOCI_MIXED = 2 };
enum { PSOCI_SIGMA_STORED = 0, PSOCI_SIGMA_DIRECT = 1 };

/* Create a context on HIP device `device`.  No reference counterpart: the
 * reference's state lives in GA-backed class instances created in the
 * drivers (driver.PSOCI.splittest.C:280-760). */
int psoci_ctx_create(int device, psoci_ctx** out);
int psoci_ctx_destroy(psoci_ctx* ctx);
/* ctx NULL: the message of the last failure on this thread of a
 * context-free call (psoci_sifs_read / psoci_sifs_write) or of a call made
 * with a NULL context ("null context"). */
const char* psoci_last_error(const psoci_ctx* ctx);

/* Integrals after the SIFS read: replaces PsociIntegrals::fetchAndProcessIntegrals
 * (PsociIntegrals.C:914-961) and its accessors fetch1eIntegralValue /
 * fetch2eIntegralValue / fetchCoreEnergy (PsociIntegrals.C:853-873, 79-82).
 * valone[NNDAF(nbf)]: h_pq at max(p,q)*(max(p,q)+1)/2 + min(p,q) (0-based p,q).
 * valtwo[NNDAF(NNDAF(nbf))]: (pq|rs) at the canonical 8-fold index of
 * PsociIntegrals::canonicalIndex (PsociIntegrals.C:892-901) minus one. */
int psoci_set_integrals(psoci_ctx* ctx, int nbf, const double* valone, const double* valtwo, double ecore);

/* SIFS-style integral files (the reference reads the Columbus moints file:
 * PsociIntegrals::fetchAndProcessIntegrals, PsociIntegrals.C:914-961, with
 * sifrh1/sifrh2/sifrd1/sifrd2).  The record layout is this library's own (the
 * Columbus library is not available as source); the header fields and the
 * processing follow the reference: 1-e records of itypea 0, itypeb != 0 are
 * summed (T + V + ...; the overlap is skipped), itypea 2 (spin-orbit x, y,
 * z) overwrite with the antisymmetric sign, 2-e values go to their canonical
 * slot (last write wins), the core energy is energy[0].  Context free; on
 * failure psoci_last_error(NULL) has the message.
 *
 * psoci_sifs_read: nbf always; valone [NNDAF(nbf)], valtwo
 * [NNDAF(NNDAF(nbf))] and ecore when not NULL (pass NULLs to query nbf).
 * psoci_sifs_write: s1 (overlap), t1, v1 (packed like valone; s1 and one of
 * t1 / v1 may be NULL), so [3][NNDAF(nbf)] or NULL (only p > q entries are
 * written), valtwo, ecore, title (NULL: a default). */
int psoci_sifs_read(const char* path, int* nbf, double* valone, double* valtwo, double* ecore);
int psoci_sifs_write(const char* path, int nbf, const double* s1, const double* t1, const double* v1,
                     const double* so, const double* valtwo, double ecore, const char* title);

/* AO-basis files for the Mulliken analysis (the reference's "aoints":
 * PsociIntegrals::fetchAndProcessOverlapIntegrals / fetchOverlapInts /
 * fetchBasisLabels, PsociIntegrals.C:90-93, 967-1110).  Same record layout;
 * psoci_sifs_read_ao sums the 1-e records with itypea = itypeb = 0 into
 * overlap [NNDAF(nbf)] (packed like valone) and copies the basis labels
 * into labels [nbf][8] (blank padded, not NUL-terminated); NULL outputs
 * are skipped (pass NULLs to query nbf).  psoci_sifs_write_ao writes the
 * overlap, the labels ([nbf][8]; NULL: "mo1", "mo2", ...) and an empty
 * 2-e part. */
int psoci_sifs_read_ao(const char* path, int* nbf, double* overlap, char* labels);
int psoci_sifs_write_ao(const char* path, int nbf, const double* overlap, const char* labels, const char* title);
/* psoci_sifs_read + psoci_set_integrals. */
int psoci_set_integrals_sifs(psoci_ctx* ctx, const char* path);

/* Configuration list in file order: replaces PsociConfigs::readConfigsToGA
 * (PsociConfigs.C:336-495) plus the determinant/SEF generation of
 * PsociDeterminants::computeConfigs (PsociDeterminants.C:1083-1191) and
 * computeNsefEntryPoints (PsociGAhamiltonian.C:4022-4048).
 * occ[ncfg][nbf] holds 0/1/2 occupations, jsym[ncfg] the spatial symmetry.
 * Each configuration gets the determinants of its jsym's class: even
 * electron counts in the state's symmetry class ((jsym + 1) / 2 ==
 * (ksym + 1) / 2: the EVEN determinants, PsociDeterminants.C:1563-1610) or
 * not (the ODD determinants of odd Ms, :1613-1658), the SEF sign following
 * the parities of jsym and ksym (PsociHamiltonian.C:1487-1495); odd
 * electron counts with one determinant per SEF (the fractional
 * determinants, :1508-1561, alpha and beta swapped when (jsym + 1) / 2 is
 * even).  Configurations of different jsym one electron apart couple
 * through the spin-orbit blocks (processSingleExcitation's SO branch and
 * socig, PsociHamiltonian.C:223-238, 709-858, with the integral from the
 * valone slot the SIFS reader's spin-orbit records overwrite); a space
 * with more than one jsym runs on the stored Hamiltonian (the direct sigma
 * and psoci_rdm return PSOCI_ERR_UNSUPPORTED for it).  nbf <= 64, at most
 * 10 open shells. */
int psoci_set_configs(psoci_ctx* ctx, int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ,
                      const int32_t* jsym);

/* Configuration files, both of the reference's readers:
 * PSOCI_CONFIGS_FILE_ORDER is PsociConfigs::readConfigsToGA (file order,
 * PsociConfigs.C:336-495, the production path); PSOCI_CONFIGS_SORTED is
 * PsociConfigs::readConfigs (std::set<pair<jsym, string>> order, duplicates
 * dropped, PsociConfigs.C:143-300).  Format: a title line, then
 * "nbf nelec ksym ncfg nopen", then one "jsym occupation" line per
 * configuration (nbf characters 0/1/2).  Each configuration line is checked
 * as the reference does: a symmetry and an occupation token, length nbf,
 * only 0/1/2, at most nopen open shells, nelec electrons; the count must
 * equal ncfg (sorted: unless duplicates were dropped).  A failure returns
 * PSOCI_ERR_ARG with *err_line the 1-based line (title = 1, header = 2; a
 * count mismatch names the last line), and psoci_last_error(NULL) the
 * message.  nopen receives the largest open-shell count found (the
 * reference adjusts the header value down to it).  Context free; NULL
 * outputs are skipped (pass occ = jsym = NULL to query ncfg and nbf first);
 * occ [ncfg][nbf] as for psoci_set_configs, jsym [ncfg], title at most
 * title_cap bytes including the terminating NUL. */
enum { PSOCI_CONFIGS_FILE_ORDER = 0, PSOCI_CONFIGS_SORTED = 1 };
int psoci_configs_read(const char* path, int order, int* nbf, int* nelec, int* ksym, int64_t* ncfg, int* nopen,
                       uint8_t* occ, int32_t* jsym, char* title, int title_cap, int64_t* err_line);
/* psoci_configs_read + psoci_set_configs (the reader's message, with its
 * line number, goes to psoci_last_error(ctx)). */
int psoci_set_configs_file(psoci_ctx* ctx, const char* path, int order);

/* Space sizes (findParams, PsociDeterminants.C:749-840): total SEFs (the H
 * dimension), total determinants, and optionally per-configuration nsef and
 * SEF offsets (arrays of length ncfg, may be NULL). */
int psoci_space_info(psoci_ctx* ctx, int64_t* nsef_total, int64_t* ndet_total, int32_t* nsef, int64_t* nsef_offset);

/* Determinants of configuration I in the reference's convention
 * (JOUTFG after computePhaseAndSort, PsociDeterminants.hpp:60-87,
 * PsociDeterminants.C:4355-4469): per determinant the alpha and beta
 * occupation bit masks (bit p = orbital p), the Ms label (n_alpha-n_beta)/2
 * and the phase bit.  Arrays of length ndet(I). */
int psoci_dets(psoci_ctx* ctx, int64_t I, uint64_t* alpha, uint64_t* beta, int32_t* ms, int32_t* phase);

/* One configuration-pair block on the GPU: PsociHamiltonian::generateSOCIblocks
 * (PsociHamiltonian.C:401-558), no threshold.  sefsef is host memory,
 * row-major nsef(I) x nsef(J); *status receives the reference's block status
 * (0 = the reference skips the block). */
int psoci_block(psoci_ctx* ctx, int64_t I, int64_t J, double* sefsef, int* status);

/* Build the stored Hamiltonian on the GPU: constructGlobalHamiltonian-
 * AltOrbMapDirectChunkJmap with packSefsef (PsociGAhamiltonian.C:2530-2857,
 * 4077-4136).  The lower triangle + diagonal (diagonal last in each row,
 * |h| >= 1e-8 or diagonal kept) is the reference's pattern; the context also
 * keeps the symmetric full matrix and an FP32 copy for sigma.  *nnz_lower
 * receives the reference's element count. */
int psoci_build_h(psoci_ctx* ctx, int64_t* nnz_lower);

/* Wall time of the last psoci_build_h on this context, in seconds:
 * seconds[0] the build proper (pair screening, element blocks, packing into
 * CSR, the diagonal and the FP32 copy, synchronised), seconds[1] the stored
 * sigma's shape autotuner that follows it.  PSOCI_ERR_STATE before a build.
 * No reference counterpart (a measurement entry point; the reference's
 * build timers are PsociGAhamiltonian.C:2822-2839). */
int psoci_build_h_timing(psoci_ctx* ctx, double* seconds);

/* Copy the stored lower triangle to host arrays: rowptr[n+1], col[nnz],
 * val[nnz], diag[n] (any may be NULL).  Replaces the GA row fetch
 * fetchLocalHblock (PsociGAhamiltonian.C:11791-11846) and fetchDIAG_SEF. */
int psoci_get_h(psoci_ctx* ctx, int64_t* rowptr, int32_t* col, double* val, double* diag);

/* Stored-Hamiltonian checkpoint: replaces the DRA dump and restart of the
 * ELL arrays cimat / icicol / number / diag (PsociConstructHamiltonian.C:
 * 56-700, driver.PSOCI.splittest.C:454-457, 690-760, 848-858,
 * --hamiltonian_restart).  The
 * file holds the lower triangle + diagonal in the reference's row layout
 * (as psoci_get_h returns it), fingerprints of the configuration list and
 * the integrals, the tuned stored-sigma shapes and a payload checksum.
 * psoci_h_load needs psoci_set_configs and psoci_set_integrals with the same
 * inputs as the writer (PSOCI_ERR_ARG otherwise, and for a truncated or
 * corrupted file; such failures leave the context as it was) and then
 * stands in for psoci_build_h: *nnz_lower as there. */
int psoci_h_save(psoci_ctx* ctx, const char* path);
int psoci_h_load(psoci_ctx* ctx, const char* path, int64_t* nnz_lower);
/* Transition Hamiltonian between the context's configuration list (rows)
 * and a second list of the same orbitals, electrons and ksym (jsym2 per
 * configuration; any mix): constructGlobalTransitionHamiltonian2WAY
 * (PsociGAhamiltonian.C:2089-2287, driver.Transition.C) with packSefsef
 * (:4077-4136).  For configuration I of the context's list and every J <= I
 * of the second (positions in the two lists) the block of the two
 * configurations, through the diagonal path when I == J whatever the
 * configurations; rows are the context's SEFs, columns J's SEFs placed at
 * the context list's entry point of J (the reference computes the entry
 * points from the first list only); |h| >= 1e-8 kept, and for I == J the
 * lower triangle with the diagonal always kept.  Refused
 * (PSOCI_ERR_UNSUPPORTED) where the reference reads out of bounds: past a
 * one-entry diagonal vnt (an equal-index pair a double excitation apart
 * whose first configuration has fewer than two open shells), or past an
 * equal-index block whose second configuration has fewer SEFs.  psoci_build_transition
 * returns the element count; psoci_get_transition copies the rows
 * (rowptr[n+1], n = the context's SEF count), columns and values to host
 * arrays (any may be NULL).  The context's stored H is left as it is. */
int psoci_build_transition(psoci_ctx* ctx, int64_t ncfg2, const uint8_t* occ2, const int32_t* jsym2, int64_t* nnz);
int psoci_get_transition(psoci_ctx* ctx, int64_t* rowptr, int32_t* col, double* val);

/* Sigma Y = H X for k vectors: matrixVectorProductsSplitChunkMultipleJchunks
 * (PsociGAhamiltonian.C:6921-7269; mode STORED) and the direct
 * ...DirectFetch path (:7549-7823; mode DIRECT, elements regenerated on the
 * fly).  X and Y are [n][k] row-major, FP64 for prec PSOCI_FP64 or FP32 for
 * PSOCI_FP32 (FP32 H values, FP32 accumulation).  on_device != 0: X and Y
 * are device pointers (no copies, launched on the context stream).  The
 * call returns once Y is written.  The context stream is non-blocking: it is
 * not ordered after the caller's streams, so the caller must finish (or
 * synchronise) the work that writes X before the call. */
int psoci_sigma(psoci_ctx* ctx, int mode, int prec, int k, const void* X, void* Y, int on_device);

/* psoci_sigma on device pointers (X, Y [n][k] row-major) that allocates
 * nothing: no reference counterpart; for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid allocation).  The zero-padded column staging of a block that
 * is not a kernel width (1, 2, 4, 8, 16, 32; 24 for the direct sigma), or
 * not 16-byte aligned, goes through a fixed 64 MiB device array per GPU
 * (held until the process ends) instead of context memory; a block needing
 * more (2 n K sizeof(element) bytes at launch width K) is refused with
 * PSOCI_ERR_ARG before anything runs.  The operator must exist already:
 * psoci_build_h (STORED) or the pair records of the direct sigma (built by
 * psoci_direct_info or a first psoci_sigma; DIRECT), else PSOCI_ERR_STATE.
 * Calls that overlap on one GPU share the array and race.  Same ordering
 * and results as psoci_sigma. */
int psoci_sigma_noalloc(psoci_ctx* ctx, int mode, int prec, int k, const void* X, void* Y);

/* Work counters of the direct sigma (built on first use): connected
 * configuration pairs visited per sigma (both orders, J = I included) and
 * structural SEF elements generated per sigma (sum of nsef_I * nsef_J over
 * them).  Replaces the reference's allcount / didblock counters
 * (PsociGAhamiltonian.C:2822-2839). */
int psoci_direct_info(psoci_ctx* ctx, int64_t* npairs, int64_t* nelems);

/* Direct sigma engine.  request: -1 leaves the choice as it is; 0 (the
 * default) the string-driven engine where the space qualifies, else the
 * configuration-driven one; 1 the configuration-driven engine (pair records
 * per connected configuration pair, elements regenerated per determinant
 * pair: the reference's scheme, PsociGAhamiltonian.C:7549-7823); 2 the
 * string-driven engine only.  The string-driven engine applies to spaces of
 * one jsym in the EVEN determinant class with an even electron count whose
 * configuration list is complete RAS: every occupation with at most m
 * electrons in a set of external orbitals (full CI, MR-CISD); it evaluates
 * the determinant Hamiltonian by resolutions of the identity over strings
 * with one and two electrons fewer, as dense products on MFMA, and agrees
 * with the configuration-driven sigma to rounding.  With engine != NULL the
 * chosen engine's operator is built (when not yet) and *engine receives 1
 * or 2; requesting 2 for a space that does not qualify fails with
 * PSOCI_ERR_UNSUPPORTED and the reason in psoci_last_error, and leaves the
 * previous choice in place.  No reference
 * counterpart (the reference has one direct scheme). */
int psoci_direct_engine(psoci_ctx* ctx, int request, int* engine);

/* The string-driven engine's plan (after psoci_direct_engine selected it):
 * external electron limit m, external orbital count, spin sectors, dense
 * products, determinants evaluated, W elements per column, and the MFMA
 * flops per column of one sigma (2 rows nq^2 summed over the products).
 * PSOCI_ERR_STATE when the engine is not set up.  Any output may be NULL. */
int psoci_direct_ras_info(psoci_ctx* ctx, int* m, int* n_external, int* nsectors, int* nblocks, int64_t* ndet,
                          int64_t* w_per_column, double* flops_per_column);

/* Times the string-driven engine: reps sigma calls of k columns in
 * precision prec on internal device buffers, with HIP events on the context
 * stream around every launch; ms[5] receives the average time per call of
 * the SEF -> determinant scatter (ms[0]), the dense products of kind
 * alpha-beta, alpha-alpha and beta-beta (ms[1..3], all blocks and column
 * chunks of the call) and the SEF gather (ms[4]); flops[3] (may be NULL) the
 * products' algorithmic flops per call by kind (2 rows nq^2 k).  No
 * reference counterpart (a measurement entry point). */
int psoci_direct_ras_timing(psoci_ctx* ctx, int prec, int k, int reps, double* ms, double* flops);

/* Connected determinant pairs one direct sigma visits: the (row
 * determinant, column determinant) pairs within two spin-orbital
 * replacements over every configuration pair the reference computes a block
 * for (J = I included) -- the work unit of the direct sigma, about 6% of
 * the structural SEF elements on an MR-CISD space.  Counted by one launch of
 * the sigma's own enumeration without the arithmetic, on first use, then
 * cached.  No reference counterpart (a measurement counter beside
 * psoci_direct_info). */
int psoci_direct_det_pairs(psoci_ctx* ctx, int64_t* det_pairs);

/* The direct sigma's launch schedule as chosen by the autotuner at pair
 * setup, one per precision: row order (0 natural, 1 by occupation string),
 * XCD run length in workgroups, and the tuner's best sample time in ms (-1
 * when fixed by the PSOCI_DIRECT_ORDER / PSOCI_DIRECT_CHUNK environment).
 * psoci_direct_schedule reports the FP32 one, psoci_direct_schedule_prec
 * the one of prec (PSOCI_FP64 or PSOCI_FP32).  No reference counterpart
 * (the reference has no launch schedule); any output may be NULL. */
int psoci_direct_schedule(psoci_ctx* ctx, int* order, int* chunk, double* tune_ms);
int psoci_direct_schedule_prec(psoci_ctx* ctx, int prec, int* order, int* chunk, double* tune_ms);

/* The stored sigma's kernel shape for precision prec (PSOCI_FP64 or
 * PSOCI_FP32) as chosen by the autotuner after psoci_build_h: lanes per
 * nonzero (1, 2, 4), rows per wavefront (1, 2, 4, 8) and the tuner's best
 * 16-column time in ms (-1 when fixed by the PSOCI_SPMM_SHAPE environment,
 * "<lanes>x<rows>").  No reference counterpart; any output may be NULL. */
int psoci_stored_shape(psoci_ctx* ctx, int prec, int* lanes_per_nz, int* rows_per_wave, double* tune_ms);

/* The stored sigma's row-length buckets (full symmetric rows of length
 * < 64, 64-255, 256-1023, 1024-4095, >= 4096): rows[5] per bucket, the
 * tuner's shape for each (lanes_per_nz[5], rows_per_wave[5]), whether the
 * per-bucket launches replace the single launch of psoci_stored_shape for
 * precision prec (*used: when they beat it by 3 % or more), and their
 * measured 16-column time (-1 when fewer than two buckets hold rows).  The
 * reference splits rows by length into its ELL table and the supplemental
 * overflow (PsociGAhamiltonian.C:4223-4312, 9325-9580); no other
 * counterpart.  Any output may be NULL. */
int psoci_stored_buckets(psoci_ctx* ctx, int prec, int* used, int64_t* rows, int* lanes_per_nz, int* rows_per_wave,
                         double* bucketed_ms);

/* Diagonal of H in the SEF basis: computeOnlyDiagElements
 * (PsociGAhamiltonian.C:13171-13250).  diag[n] host. */
int psoci_diag(psoci_ctx* ctx, double* diag);

typedef struct {
  int nroots;         /* k: lowest roots sought */
  int prec;           /* PSOCI_FP64, PSOCI_FP32 (sigma only) or PSOCI_MIXED */
  int sigma_mode;     /* PSOCI_SIGMA_STORED or PSOCI_SIGMA_DIRECT */
  int max_iter;       /* block iterations */
  int max_subspace;   /* basis size cap before collapse (0: 5*nroots) */
  int guess;          /* 0: unit vectors on the k lowest diagonal elements;
                         1: unit vectors on the first k SEFs (PsociGAvectors.C:687-726);
                         2: the caller's vectors, passed in evecs ([n][k] row-major, any
                            linearly independent set; orthonormalised here), as a restart
                            from stored CI vectors (PsociGAvectors.C:539-659) */
  double tol;         /* residual-norm tolerance for convergence */
  double switch_tol;  /* PSOCI_MIXED: FP32 sigma until max residual < switch_tol */
  int extra_roots;    /* guard roots carried in the block beyond nroots (convergence is
                         tested on the nroots lowest only); with PSOCI_MIXED they ride in
                         the FP32 phase only and the FP64 phase continues with nroots;
                         0 for the plain solve */
  int precond;        /* 0: the reference's diagonal preconditioner r_i / (D_ii - lambda)
                            (PsociGAbasis.C:1860-1920); 1: the same with Olsen's correction
                            against the Ritz vector x: t = f (r - eps x),
                            eps = (x . f r) / (x . f x).  Both skip |D_ii - lambda| <= 1e-3 */
  int graph;          /* 0: with the stored sigma, each step's device work (sigma of the new
                            block and its projection V^T H V) runs as a captured hipGraph per
                            block shape, replayed; 1: always launched eagerly.  The results
                            are the same bit for bit */
  int select_rule;    /* correction vectors added per iteration.  1: the reference's rule,
                            one per root whose residual is not yet below the tolerance
                            (PsociGAbasis.C:1371-1470); 0: the same, and the columns by
                            which the sigma launch pads that block to its kernel width
                            (1, 2, 4, 8, 16, 32) take the corrections of unconverged guard
                            roots, then of converged wanted roots with the largest
                            residuals (above 1e-3 x the tolerance): they cost no sigma
                            time and cut iterations */
} psoci_dav_params;

typedef struct {
  int iterations;     /* block iterations */
  int sigma_calls;    /* sigma launches (each covers a block of vectors) */
  int sigma_vectors;  /* vectors multiplied */
  int switch_iter;    /* iteration of the FP32->FP64 switch (-1: none) */
  int converged;
  double seconds;     /* wall time of the solve */
  double sigma_seconds;
  double max_rnorm;
  double eig_seconds;  /* host time in the subspace eigensolves (Rayleigh-Ritz, orthonormalisation) */
  int graph_launches;  /* captured steps replayed */
} psoci_dav_stats;

/* Block Davidson on the GPU: replaces the production driver loop
 * (driver.PSOCI.splittest.C:895-1380) with PsociGAbasis / PsociGAsubspace
 * (subspace VtV/VtHV, dsygv, residuals, diagonal preconditioner with
 * DAVIDSON_MIN = 1e-3, Ritz collapse).  evals[k] (without ecore), evecs
 * [n][k] row-major host (may be NULL), stats may be NULL.  Device memory
 * for the basis V and HV, n x max_subspace each: FP64 with PSOCI_FP64, FP32
 * with PSOCI_FP32, both with PSOCI_MIXED (its FP32 phase keeps an FP32
 * basis); plus FP64 work blocks of n x k and n x 2k. */
int psoci_davidson(psoci_ctx* ctx, const psoci_dav_params* params, double* evals, double* evecs,
                   psoci_dav_stats* stats);

/* CI-vector restart files in the reference's binary vector format
 * (PsociVector.C:530-553, 627-643, 685-693; written by
 * PsociGAvectors::dumpVectorScalable, PsociGAvectors.C:207-325, at the end
 * of the driver, driver.PSOCI.splittest.C:1354-1358; read by
 * readVectorScalable + specifyDiagonalGuessVectors, PsociGAvectors.C:539-659,
 * 687-726, for --vector_restart, driver.PSOCI.splittest.C:962-971):
 * int64 nbasis, int64 nroots, int32 width, int32 prec, then per root an
 * int64 root number and nbasis doubles, then the title's bytes.  Context
 * free; on failure psoci_last_error(NULL) has the message.
 *
 * psoci_civec_write: C [n][k] row-major (k roots), title NULL: "default".
 * psoci_civec_info: nbasis, nroots and the title (NUL-terminated, at most
 * title_cap - 1 bytes) of a file; any output may be NULL.
 * psoci_civec_read: n must equal the file's nbasis; fills C [n][k] with the
 * first min(nroots, k) roots and completes columns j >= *nread with unit
 * vectors e_j (requires k <= n).  The result is a guess for psoci_davidson
 * with guess = 2. */
int psoci_civec_write(const char* path, int64_t n, int k, const double* C, const char* title);
int psoci_civec_info(const char* path, int64_t* nbasis, int64_t* nroots, char* title, int64_t title_cap);
int psoci_civec_read(const char* path, int64_t n, int k, double* C, int* nread);
/* Subspace projection out = A^T B over n rows on FP64 MFMA: the VtV / VtHV
 * dot products of PsociGAsubspace::generateIncrementalSubspace
 * (PsociGAsubspace.C:273-362) and the MGS overlaps of
 * PsociGAbasis::orthonormalizeAppendedVectors (PsociGAbasis.C:1371-1470) as
 * one product.  A [n][lda] and B [n][ldb] are device pointers (row-major,
 * first m resp. b columns used); out is host memory, m x b row-major. */
int psoci_project(psoci_ctx* ctx, int64_t n, int m, int b, const double* A, int lda, const double* B, int ldb,
                  double* out);
/* Basis update Out = alpha Out + A Y over n rows on FP64 MFMA: the Ritz
 * collapse of PsociGAbasis::expandAndReplaceVectorsNew
 * (PsociGAbasis.C:483-579) and the residual combinations of
 * PsociGAsubspace::computeRnormSingle (PsociGAsubspace.C:874-989) as one
 * product.  A [n][lda] (first m columns) and Out [n][ldo] (first k columns)
 * are device pointers, row-major, not overlapping; Y is host memory, m x k
 * row-major; alpha 0 ignores Out's old contents.  Returns once Out is
 * written. */
int psoci_combine(psoci_ctx* ctx, int64_t n, int m, int k, const double* A, int lda, const double* Y, double alpha,
                  double* Out, int ldo);

/* Adjoint of E_k: dE/dtheta = c^T (dH/dtheta) c for every packed integral,
 * accumulated over k vectors C[n][k] (row-major, host) with weights w[k]
 * (NULL: all 1).  grad_one[NNDAF(nbf)], grad_two[NNDAF(NNDAF(nbf))];
 * grad_so[NNDAF(nbf)] is the spin-orbit gradient (identically zero: the
 * adjoint takes spaces of one jsym, where the reference never enters socig,
 * PsociHamiltonian.C:223-238; others return PSOCI_ERR_UNSUPPORTED).
 * No reference counterpart for the 2-RDM; the 1-RDM part relates to
 * PsociNOpopulations::constructDensityMatrix (PsociNOpopulations.C:858-989)
 * as grad_one[pp] = D_pp and grad_one[pq] = 2 D_pq (p != q). */
int psoci_rdm(psoci_ctx* ctx, int k, const double* C, const double* w, double* grad_one, double* grad_two,
              double* grad_so);

/* CI energy analysis: analyzeEnergyContributions (PsociCIAnalysis.C:42-267,
 * the per-configuration sums of :126-142) on the GPU.  For k roots C [n][k]
 * (row-major, host) with eigenvalues evals[k] (without ecore), deltae
 * [ncfg][k] receives, per configuration I and root r, the sum over the SEFs i
 * of I of -c_ir^2 (H_ii - E_r) / (1 - c_ir^2) in the reference's order and
 * rounding (an SEF with c_ir^2 == 1 resets the sum to E_r).  Sorting by
 * |deltae| and the threshold of the report are the caller's (psoci_amd.api
 * analyze_energy_contributions mirrors the reference's report). */
int psoci_energy_contrib(psoci_ctx* ctx, int k, const double* C, const double* evals, double* deltae);
/* Natural orbitals from a CI one-particle density (context free, host):
 * constructNaturalOrbitals with diagonalizeNOinMObasis, sortNOinMObasis,
 * transformNOstoAObasis and computeMOdensity (PsociNOpopulations.C:173-500,
 * 569-729).  rdm1 [NNDAF(nbf)] is the reference's packed CI density (the
 * off-diagonal counted once per pair: psoci_rdm's grad_one with the
 * off-diagonal halved).  occ [nbf]: occupations, largest first; no_mo
 * [nbf][nbf]: row i = natural orbital i over the MOs (sign: largest-magnitude
 * coefficient positive).  With mocoef [nbf MOs][nbf AOs] (else NULL):
 * no_ao [nbf][nbf] (row i over the AOs) and dens_ao [NNDAF(nbf)], the AO
 * density C^T D C packed.  Any output may be NULL. */
int psoci_natural_orbitals(int nbf, const double* rdm1, const double* mocoef, double* occ, double* no_mo,
                           double* no_ao, double* dens_ao);
/* Mulliken analysis (context free, host): constructMullikenPopulationDensAO
 * (PsociNOpopulations.C:1084-1206).  population [nbf][nbf] = P_{mu nu}
 * S_{mu nu} from the packed AO density and overlap; *total = its sum (the
 * electron count for an S-orthonormal MO set). */
int psoci_mulliken(int nbf, const double* dens_ao, const double* overlap, double* population, double* total);
/* Basis-function groups for condensed populations: clusterBasisLabels
 * (PsociNOpopulations.C:1451-1510).  labels [nbf][8]; mode 0 atomic (leading
 * digits and non-alphanumerics dropped, cut at the first '_'), 1 subshell
 * (leading digits and non-alphanumerics dropped), 2 the labels as they are.
 * group [nbf] receives the group of each function, numbered in the sorted
 * order of the condensed labels; *ngroups their count. */
int psoci_basis_groups(int nbf, const char* labels, int mode, int32_t* group, int32_t* ngroups);
/* Condensed Mulliken populations: atomicPopulations / subshellPopulations
 * (gross [ngroups]: sum over the functions of a group of their population
 * columns) and atomicOverlapPopulations (bond [ngroups][ngroups], g <= h:
 * the block sum, doubled for g != h; entries below the diagonal are 0)
 * (PsociNOpopulations.C:1513-1790), summed in the reference's order.  Either
 * output may be NULL. */
int psoci_group_populations(int nbf, const double* population, const int32_t* group, int ngroups, double* gross,
                            double* bond);
/* Device handles of the context (for zero-copy callers and for timing with
 * HIP events on the stream the library launches on): which = 0: diag
 * (double[n]); which = 1: the context's hipStream_t. */
int psoci_device_ptr(psoci_ctx* ctx, int which, void** ptr);

/* Synchronise the context stream. */
int psoci_sync(psoci_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* PSOCI_H */
