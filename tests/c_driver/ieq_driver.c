/* Plain C driver of libinductioneq_hip.so (no Python, no torch): the flow of
 * matlab/+induction_eq/compute_numerical_solution.m:23-41 through the C ABI.
 *
 *   ieq_driver <settings> <module>... -- <integrator> <nsteps> <in.bin> <out.bin>
 *
 * Builds the program from the generate_settings string and module list, uploads the
 * AoS REAL4 field of in.bin (num_nodes_pad x 4 REALs), runs nsteps steps of the
 * integrator four times: through ieq_step, through ieq_run with the reference kernel
 * list once per step, through the typed per-kernel entry points (inductioneq_kernels.h)
 * with the same list, and through one ieq_run on the list repeated nsteps times (the
 * reference's repmat(RK_Step, 1, nsteps), compute_numerical_solution.m:91-108); writes the
 * four results (AoS REAL4, in that order) to out.bin and prints the path each one took
 * (ieq_last_step_path: 1 fused kernels, 0 reference kernels). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/inductioneq.h"
#include "../../include/inductioneq_kernels.h"

static int fail(ieq_program *p, const char *what) {
  fprintf(stderr, "ieq_driver: %s: %s\n", what, ieq_last_error(p));
  return 1;
}

/* RK_Step kernel list of CarpenterKennedy2N54 with the Hall term
 * (compute_numerical_solution_setup.m:319-340) */
static const char *ck54[] = {
    "calc_curlB_rho_1_2args", "CarpenterKennedy2N54_1a", "CarpenterKennedy2N54_1b",
    "calc_curlB_rho_1_2args", "CarpenterKennedy2N54_2a", "CarpenterKennedy2N54_2b",
    "calc_curlB_rho_1_2args", "CarpenterKennedy2N54_3a", "CarpenterKennedy2N54_3b",
    "calc_curlB_rho_1_2args", "CarpenterKennedy2N54_4a", "CarpenterKennedy2N54_4b",
    "calc_curlB_rho_1_2args", "CarpenterKennedy2N54_5a", "CarpenterKennedy2N54_5b", "calc_time_2_args"};

/* The same list as typed calls (b1, b2, curlB_rho, u, rho, time). */
typedef int (*ieq_kernel6)(ieq_program *, const uint32_t *, const uint32_t *, void *, void *, void *, void *, void *,
                           void *, double *);
static const ieq_kernel6 ck54_typed[] = {
    ieq_calc_curlB_rho_1_2args, ieq_CarpenterKennedy2N54_1a, ieq_CarpenterKennedy2N54_1b,
    ieq_calc_curlB_rho_1_2args, ieq_CarpenterKennedy2N54_2a, ieq_CarpenterKennedy2N54_2b,
    ieq_calc_curlB_rho_1_2args, ieq_CarpenterKennedy2N54_3a, ieq_CarpenterKennedy2N54_3b,
    ieq_calc_curlB_rho_1_2args, ieq_CarpenterKennedy2N54_4a, ieq_CarpenterKennedy2N54_4b,
    ieq_calc_curlB_rho_1_2args, ieq_CarpenterKennedy2N54_5a, ieq_CarpenterKennedy2N54_5b, ieq_calc_time_2_args};

int main(int argc, char **argv) {
  int sep = 0;
  for (int i = 2; i < argc; ++i)
    if (!strcmp(argv[i], "--")) sep = i;
  if (argc < 4 || !sep || argc != sep + 5) {
    fprintf(stderr, "usage: ieq_driver <settings> <module>... -- CarpenterKennedy2N54 <nsteps> <in.bin> <out.bin>\n");
    return 2;
  }
  const char *settings = argv[1];
  const char *const *modules = (const char *const *)&argv[2];
  const int nmod = sep - 2;
  const char *integrator = argv[sep + 1];
  const long long nsteps = atoll(argv[sep + 2]);
  if (strcmp(integrator, "CarpenterKennedy2N54")) {
    fprintf(stderr, "ieq_driver: this driver carries the kernel list of CarpenterKennedy2N54 only\n");
    return 2;
  }
  ieq_program *p = NULL;
  if (ieq_program_build(&p, settings, modules, nmod) != IEQ_OK) return fail(NULL, "build");
  ieq_program_info info;
  if (ieq_program_get_info(p, &info) != IEQ_OK) return fail(p, "info");
  const size_t n = (size_t)info.num_nodes_pad, rb = (size_t)info.real_bytes;
  const int npaths = 4;
  void *host = malloc(n * 4 * rb), *out = malloc(npaths * n * 4 * rb);
  FILE *f = fopen(argv[sep + 3], "rb");
  if (!f || fread(host, rb, n * 4, f) != n * 4) {
    fprintf(stderr, "ieq_driver: cannot read %s\n", argv[sep + 3]);
    return 1;
  }
  fclose(f);
  printf("ieq_driver: paths");
  for (int path = 0; path < npaths; ++path) {
    /* b1, b2, curlB_rho, u, rho, time (kernel_time_integrator.cl:1023-1026) */
    void *b1, *b2, *J, *u, *rho, *t;
    const size_t vb = ieq_vector_field_bytes(p), sb = ieq_scalar_field_bytes(p);
    if (ieq_alloc(p, vb, &b1) || ieq_alloc(p, vb, &b2) || ieq_alloc(p, vb, &J) || ieq_alloc(p, vb, &u) ||
        ieq_alloc(p, sb, &rho) || ieq_alloc(p, 2 * rb, &t))
      return fail(p, "alloc");
    if (ieq_upload_real4(p, b1, host, n) != IEQ_OK) return fail(p, "upload");
    /* rho = 1 everywhere */
    void *ones = malloc(n * rb);
    for (size_t i = 0; i < n; ++i) {
      if (rb == 8) ((double *)ones)[i] = 1.0;
      else ((float *)ones)[i] = 1.0f;
    }
    if (ieq_upload_real(p, rho, ones, n) != IEQ_OK) return fail(p, "upload rho");
    free(ones);
    void *args[6] = {b1, b2, J, u, rho, t};
    const int nk = (int)(sizeof(ck54) / sizeof(ck54[0]));
    if (path == 0) {
      if (ieq_step(p, integrator, nsteps, 0, args, 6, NULL) != IEQ_OK) return fail(p, "ieq_step");
    } else if (path == 3) {
      const uint32_t g[3] = {(uint32_t)info.nx, (uint32_t)info.ny, (uint32_t)info.nz}, l[3] = {(uint32_t)info.w_size, 1, 1};
      const char **list = (const char **)malloc((size_t)nsteps * nk * sizeof(char *));
      for (long long s = 0; s < nsteps; ++s)
        for (int k = 0; k < nk; ++k) list[s * nk + k] = ck54[k];
      double sec = 0;
      if (ieq_run(p, list, (int)(nsteps * nk), g, l, args, 6, &sec) != IEQ_OK) return fail(p, "ieq_run (repeated list)");
      free(list);
    } else {
      const uint32_t g[3] = {(uint32_t)info.nx, (uint32_t)info.ny, (uint32_t)info.nz}, l[3] = {(uint32_t)info.w_size, 1, 1};
      for (long long s = 0; s < nsteps; ++s) {
        if (path == 1) {
          if (ieq_run(p, ck54, nk, g, l, args, 6, NULL) != IEQ_OK) return fail(p, "ieq_run");
        } else {
          for (size_t k = 0; k < sizeof(ck54_typed) / sizeof(ck54_typed[0]); ++k)
            if (ck54_typed[k](p, g, l, b1, b2, J, u, rho, t, NULL) != IEQ_OK) return fail(p, ck54[k]);
        }
      }
    }
    printf(" %d", ieq_last_step_path(p));
    if (ieq_download_real4(p, (char *)out + path * n * 4 * rb, b1, n) != IEQ_OK) return fail(p, "download");
    void *all[6] = {b1, b2, J, u, rho, t};
    for (int k = 0; k < 6; ++k) ieq_free(p, all[k]);
  }
  printf("\n");
  f = fopen(argv[sep + 4], "wb");
  if (!f || fwrite(out, rb, npaths * n * 4, f) != npaths * n * 4) {
    fprintf(stderr, "ieq_driver: cannot write %s\n", argv[sep + 4]);
    return 1;
  }
  fclose(f);
  free(host);
  free(out);
  ieq_program_release(p);
  printf("ieq_driver: %lld step(s) of %s on %dx%dx%d done\n", nsteps, integrator, info.nx, info.ny, info.nz);
  return 0;
}
