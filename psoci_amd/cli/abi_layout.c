/* Prints sizeof and offsetof of the ABI's structs as the C compiler lays
 * them out, one "name value" per line, for the ctypes mirror's check.  No
 * device calls. */
#include <stddef.h>
#include <stdio.h>

#include "psoci.h"

#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))

int main(void) {
  printf("psoci_dav_params %zu\n", sizeof(psoci_dav_params));
  F(psoci_dav_params, nroots);
  F(psoci_dav_params, prec);
  F(psoci_dav_params, sigma_mode);
  F(psoci_dav_params, max_iter);
  F(psoci_dav_params, max_subspace);
  F(psoci_dav_params, guess);
  F(psoci_dav_params, tol);
  F(psoci_dav_params, switch_tol);
  F(psoci_dav_params, extra_roots);
  F(psoci_dav_params, precond);
  F(psoci_dav_params, graph);
  F(psoci_dav_params, select_rule);
  printf("psoci_dav_stats %zu\n", sizeof(psoci_dav_stats));
  F(psoci_dav_stats, iterations);
  F(psoci_dav_stats, sigma_calls);
  F(psoci_dav_stats, sigma_vectors);
  F(psoci_dav_stats, switch_iter);
  F(psoci_dav_stats, converged);
  F(psoci_dav_stats, seconds);
  F(psoci_dav_stats, sigma_seconds);
  F(psoci_dav_stats, max_rnorm);
  F(psoci_dav_stats, eig_seconds);
  F(psoci_dav_stats, graph_launches);
  return 0;
}
