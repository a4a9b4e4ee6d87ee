/* psoci_run: command-line driver over the C ABI (include/psoci.h), in C99.
 *
 * Stands in for the reference's production driver (driver.PSOCI.splittest.C:
 * option parsing parseUserValues :69-270, main :280, the Davidson run
 * :895-1340 and its report :1354-1380) on one GPU: configuration file +
 * SIFS-style integral file in, the lowest roots out, printed as the
 * reference prints them ("Eigenvalue i Total energy (au) ... Electronic
 * (au) ... rnorm ...").  Options keep the reference's letters where the
 * meaning carries over:
 *   -c FILE  configuration file (configs_filename)
 *   -m FILE  SIFS-style integral file (moints_filename)
 *   -n K     roots (numRoots)
 *   -v FILE  CI-vector file written at the end (vector_filename)
 *   -r       start from the vectors in the -v file (vector_restart)
 *   -H       load the stored Hamiltonian from the --hamiltonian-file
 *            checkpoint instead of building it (hamiltonian_restart)
 *   -W       write that checkpoint after the build (hamiltonian_store)
 *   -x MULT  subspace cap = MULT * K vectors (subspace_max_mult)
 *   -I N     maximum iterations (subspace_iterations)
 *   -F TOL   residual-norm tolerance (final_rtol)
 *   -S       also print the bare eigenvalues (showevals)
 *   -V, -h   version, help
 * and this library's own:
 *   --direct            direct sigma (no stored Hamiltonian)
 *   --prec fp64|fp32|mixed, --switch TOL, --guard N,
 *   --precond 0|1       0 the reference's diagonal preconditioner, 1 with Olsen's correction
 *   --sorted            read the configuration file sorted and de-duplicated
 *   --hamiltonian-file FILE (default psoci_hamiltonian.bin)
 * Exit status: 0 converged, 2 not converged, 1 error. */
#define _GNU_SOURCE
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "psoci.h"

static int check(psoci_ctx* ctx, int rc, const char* what) {
  if (rc != PSOCI_OK) {
    fprintf(stderr, "psoci_run: %s failed (%d): %s\n", what, rc, psoci_last_error(ctx));
    return 0;
  }
  return 1;
}

static void usage(void) {
  fprintf(stderr,
          "usage: psoci_run -c configs -m moints [-n roots] [-F tol] [-I iters] [-x mult] [-v vecfile [-r]]\n"
          "                 [-W|-H] [--hamiltonian-file F] [--direct] [--prec fp64|fp32|mixed] [--switch tol]\n"
          "                 [--guard n] [--precond 0|1] [--sorted] [-S] [-V] [-h]\n");
}

int main(int argc, char** argv) {
  const char* configs = NULL;
  const char* moints = NULL;
  const char* vecfile = NULL;
  const char* hfile = "psoci_hamiltonian.bin";
  int nroots = 1, restart = 0, hload = 0, hstore = 0, maxmult = 0, iters = 0, show = 0, sorted = 0;
  int mode = PSOCI_SIGMA_STORED, prec = PSOCI_FP64, guard = 0, precond = 0;
  double tol = 1e-6, switch_tol = 1e-4;
  static const struct option longopts[] = {
      {"direct", no_argument, 0, 1},          {"prec", required_argument, 0, 2},
      {"switch", required_argument, 0, 3},    {"guard", required_argument, 0, 4},
      {"precond", required_argument, 0, 5},   {"sorted", no_argument, 0, 6},
      {"hamiltonian-file", required_argument, 0, 7}, {"help", no_argument, 0, 'h'},
      {0, 0, 0, 0}};
  int opt;
  while ((opt = getopt_long(argc, argv, "c:m:n:v:rHWx:I:F:SVh", longopts, NULL)) != -1) {
    switch (opt) {
      case 'c': configs = optarg; break;
      case 'm': moints = optarg; break;
      case 'n': nroots = atoi(optarg); break;
      case 'v': vecfile = optarg; break;
      case 'r': restart = 1; break;
      case 'H': hload = 1; break;
      case 'W': hstore = 1; break;
      case 'x': maxmult = atoi(optarg); break;
      case 'I': iters = atoi(optarg); break;
      case 'F': tol = atof(optarg); break;
      case 'S': show = 1; break;
      case 'V': printf("psoci_run (PSOCI determinant Hamiltonian and Davidson on one MI355X)\n"); return 0;
      case 1: mode = PSOCI_SIGMA_DIRECT; break;
      case 2:
        if (!strcmp(optarg, "fp64")) prec = PSOCI_FP64;
        else if (!strcmp(optarg, "fp32")) prec = PSOCI_FP32;
        else if (!strcmp(optarg, "mixed")) prec = PSOCI_MIXED;
        else { usage(); return 1; }
        break;
      case 3: switch_tol = atof(optarg); break;
      case 4: guard = atoi(optarg); break;
      case 5: precond = atoi(optarg); break;
      case 6: sorted = 1; break;
      case 7: hfile = optarg; break;
      case 'h': usage(); return 0;
      default: usage(); return 1;
    }
  }
  if (!configs || !moints || nroots < 1 || (restart && !vecfile) || (hload && hstore)) {
    usage();
    return 1;
  }
  psoci_ctx* ctx = NULL;
  if (psoci_ctx_create(0, &ctx) != PSOCI_OK) {
    fprintf(stderr, "psoci_run: no HIP device\n");
    return 1;
  }
  int status = 1;
  double* evals = NULL;
  double* evecs = NULL;
  int64_t n = 0, ndet = 0, nnz = 0;
  double ecore = 0.0;
  int nbf = 0;
  if (!check(ctx, psoci_set_configs_file(ctx, configs, sorted ? PSOCI_CONFIGS_SORTED : PSOCI_CONFIGS_FILE_ORDER),
             "reading the configuration file"))
    goto done;
  if (!check(ctx, psoci_set_integrals_sifs(ctx, moints), "reading the integral file")) goto done;
  if (psoci_sifs_read(moints, &nbf, NULL, NULL, &ecore) != PSOCI_OK) {
    fprintf(stderr, "psoci_run: %s\n", psoci_last_error(NULL));
    goto done;
  }
  if (!check(ctx, psoci_space_info(ctx, &n, &ndet, NULL, NULL), "space")) goto done;
  printf("configurations read; %lld SEFs, %lld determinants, %d orbitals, core energy %.12f\n", (long long)n,
         (long long)ndet, nbf, ecore);
  if (mode == PSOCI_SIGMA_STORED) {
    if (hload) {
      if (!check(ctx, psoci_h_load(ctx, hfile, &nnz), "loading the Hamiltonian checkpoint")) goto done;
      printf("Hamiltonian loaded from %s: %lld stored elements\n", hfile, (long long)nnz);
    } else {
      if (!check(ctx, psoci_build_h(ctx, &nnz), "building the Hamiltonian")) goto done;
      printf("Hamiltonian built: %lld stored elements (lower triangle and diagonal)\n", (long long)nnz);
      if (hstore) {
        if (!check(ctx, psoci_h_save(ctx, hfile), "writing the Hamiltonian checkpoint")) goto done;
        printf("Hamiltonian written to %s\n", hfile);
      }
    }
  }
  if ((int64_t)nroots > n) nroots = (int)n;
  evals = (double*)calloc((size_t)nroots, sizeof(double));
  evecs = (double*)calloc((size_t)n * (size_t)nroots, sizeof(double));
  if (!evals || !evecs) {
    fprintf(stderr, "psoci_run: out of host memory\n");
    goto done;
  }
  {
    psoci_dav_params p;
    psoci_dav_stats st;
    int nread = 0, rc;
    memset(&p, 0, sizeof p);
    memset(&st, 0, sizeof st);
    p.nroots = nroots;
    p.prec = prec;
    p.sigma_mode = mode;
    p.max_iter = iters;
    p.max_subspace = maxmult > 0 ? maxmult * nroots : 0;
    p.guess = 0;
    p.tol = tol;
    p.switch_tol = switch_tol;
    p.extra_roots = guard;
    p.precond = precond;
    if (restart) {
      if (!check(ctx, psoci_civec_read(vecfile, n, nroots, evecs, &nread), "reading the vector file")) goto done;
      printf("restart: %d vectors read from %s\n", nread, vecfile);
      p.guess = 2;
      p.extra_roots = 0;
    }
    rc = psoci_davidson(ctx, &p, evals, evecs, &st);
    if (rc != PSOCI_OK && rc != PSOCI_ERR_CONVERGE) {
      check(ctx, rc, "Davidson");
      goto done;
    }
    if (show)
      for (int i = 0; i < nroots; ++i) printf("root %d %.12f\n", i, evals[i]);
    printf("Number of total iterations is %d\n", st.iterations);
    printf("TOTAL MATRIX VECTOR PRODUCTS is %d\n", st.sigma_vectors);
    printf("Current results: fcore is %.12f\n", ecore);
    for (int i = 0; i < nroots; ++i)
      printf("Eigenvalue %d Total energy (au) %.12f Electronic (au) %.12f\n", i, evals[i] + ecore, evals[i]);
    printf("max rnorm %.3e; all converged = %d; %.3f s\n", st.max_rnorm, st.converged, st.seconds);
    if (vecfile && !check(ctx, psoci_civec_write(vecfile, n, nroots, evecs, "psoci_run"), "writing the vector file"))
      goto done;
    status = rc == PSOCI_OK ? 0 : 2;
  }
done:
  free(evals);
  free(evecs);
  psoci_ctx_destroy(ctx);
  return status;
}
