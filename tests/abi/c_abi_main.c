/*
 * A C caller of the reference's drop-in C ABI, compiled with gcc against a
 * generated mkpp.h and linked to libmkpp_<mechanism>.so.  Buffers are sized
 * with the header's MKPP_NUM_* macros, as a host model sizes them; the solver
 * options are the drop-in's generation-time defaults (no mkpp_amd_* call).
 *
 * usage: c_abi_main IN OUT DT
 *   IN:  int64 num_cells, then float64 conc[MKPP_NUM_SPECIES * num_cells]
 *        (LayoutLeft), temp[num_cells], pressure[num_cells],
 *        air_density[num_cells], params[MKPP_NUM_RATE_PARAMETERS * num_cells]
 *   OUT: int32 return code of mkpp_integrate, then the integrated conc
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "mkpp.h"

static int read_all(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s IN OUT DT\n", argv[0]);
    return 2;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) return 2;
  int64_t n = 0;
  if (!read_all(in, &n, sizeof n) || n <= 0) return 2;
  const size_t cells = (size_t)n;
  double* conc = malloc(sizeof(double) * MKPP_NUM_SPECIES * cells);
  double* temp = malloc(sizeof(double) * cells);
  double* pres = malloc(sizeof(double) * cells);
  double* rho = malloc(sizeof(double) * cells);
  double* params = malloc(sizeof(double) * (MKPP_NUM_RATE_PARAMETERS > 0 ? MKPP_NUM_RATE_PARAMETERS : 1) * cells);
  if (!read_all(in, conc, sizeof(double) * MKPP_NUM_SPECIES * cells) || !read_all(in, temp, sizeof(double) * cells) ||
      !read_all(in, pres, sizeof(double) * cells) || !read_all(in, rho, sizeof(double) * cells) ||
      !read_all(in, params, sizeof(double) * MKPP_NUM_RATE_PARAMETERS * cells))
    return 2;
  fclose(in);

  mkpp_handle_t h = NULL;
  size_t ns = 0;
  int rc = mkpp_create_handle(&h, cells);
  if (rc == MKPP_SUCCESS) rc = mkpp_initialize(h);
  if (rc == MKPP_SUCCESS) rc = mkpp_get_species_count(h, &ns);
  if (rc == MKPP_SUCCESS && ns != MKPP_NUM_SPECIES) rc = -1;
  if (rc == MKPP_SUCCESS) rc = mkpp_set_state_ptrs(h, conc, temp, pres, rho);
  if (rc == MKPP_SUCCESS && MKPP_NUM_RATE_PARAMETERS > 0)
    rc = mkpp_set_rate_parameter_ptrs(h, params, MKPP_NUM_RATE_PARAMETERS);
  if (rc == MKPP_SUCCESS) rc = mkpp_integrate(h, atof(argv[3]));
  if (rc != MKPP_SUCCESS && h) {
    char msg[256];
    mkpp_get_last_error(h, msg, sizeof msg);
    fprintf(stderr, "mkpp error %d: %s\n", rc, msg);
  }
  if (h) mkpp_destroy_handle(h);

  FILE* out = fopen(argv[2], "wb");
  if (!out) return 2;
  int32_t rc32 = rc;
  fwrite(&rc32, sizeof rc32, 1, out);
  fwrite(conc, sizeof(double), MKPP_NUM_SPECIES * cells, out);
  fclose(out);
  printf("c_abi_main: %lld cells of %d species, %d rate parameters, method %d, rc %d\n", (long long)n,
         MKPP_NUM_SPECIES, MKPP_NUM_RATE_PARAMETERS, MKPP_DEFAULT_METHOD, rc);
  free(conc);
  free(temp);
  free(pres);
  free(rho);
  free(params);
  return rc == MKPP_SUCCESS ? 0 : 1;
}
