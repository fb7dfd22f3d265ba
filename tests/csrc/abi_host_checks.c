/* Host-side checks of the C ABI, run under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_cpu_sanitizers.py builds the C-ABI
 * unit with -fsanitize=address,undefined for the host and links it with this
 * driver).  Needs no GPU: it drives every entry point through its argument
 * checks (null plan, null pointers, out-of-range sizes), and plan creation
 * over a sweep of grid lengths, precisions and full/half grids through the
 * length factorisation, the LDS-size checks, the lattice inversion and --
 * without a GPU -- the clean-up of a plan whose device allocation failed.
 * Exit status 0 and a last line "ok <n> checks" on success. */
#include <stdio.h>
#include <string.h>

#include "ofdft_amd.h"

static int nchecks = 0, nfail = 0;

static void expect_error(int rc, const char* what) {
  ++nchecks;
  const char* msg = ofdft_last_error();
  if (rc >= 0 || !msg || !msg[0]) {
    ++nfail;
    fprintf(stderr, "FAIL %s: rc=%d msg='%s'\n", what, rc, msg ? msg : "(null)");
  }
}

static void expect_code(int rc, int code, const char* what) {
  ++nchecks;
  if (rc != code) {
    ++nfail;
    fprintf(stderr, "FAIL %s: rc=%d, expected %d (%s)\n", what, rc, code, ofdft_last_error());
  }
}

static void expect_msg(const char* needle, const char* what) {
  ++nchecks;
  const char* msg = ofdft_last_error();
  if (!msg || !strstr(msg, needle)) {
    ++nfail;
    fprintf(stderr, "FAIL %s: message '%s' lacks '%s'\n", what, msg ? msg : "(null)", needle);
  }
}

/* 2^a 3^b 5^c 7^d 11^e 13^f */
static int smooth13(int n) {
  const int p[6] = {2, 3, 5, 7, 11, 13};
  for (int i = 0; i < 6; ++i)
    while (n % p[i] == 0) n /= p[i];
  return n == 1;
}

int main(void) {
  const ofdft_plan P0 = NULL;
  double lat[9] = {10, 0, 0, 0, 11, 0, 0, 0, 12};
  double sing[9] = {1, 0, 0, 2, 0, 0, 0, 0, 1};
  double d[16];
  long long ll[16];
  int iv[16];
  const void* vp[2] = {NULL, NULL};
  void* wp[2] = {NULL, NULL};

  ++nchecks;
  if (ofdft_version() != 1) {
    ++nfail;
    fprintf(stderr, "FAIL version\n");
  }

  /* every entry point with a null plan */
  expect_code(ofdft_plan_destroy(P0), 0, "destroy(NULL) is a no-op");
  expect_code(ofdft_plan_set_graph(P0, 1), OFDFT_EINVAL, "set_graph");
  expect_code(ofdft_plan_set_tiles(P0, 16, 16), OFDFT_EINVAL, "set_tiles");
  expect_code(ofdft_plan_set_stream(P0, NULL), OFDFT_EINVAL, "set_stream");
  expect_code(ofdft_plan_set_profiling(P0, 1), OFDFT_EINVAL, "set_profiling");
  expect_code(ofdft_plan_get_profile(P0, d, ll, d), OFDFT_EINVAL, "get_profile");
  expect_code(ofdft_set_functional(P0, NULL), OFDFT_EINVAL, "set_functional");
  expect_code(ofdft_set_vloc(P0, NULL), OFDFT_EINVAL, "set_vloc");
  expect_code(ofdft_eval(P0, NULL, NULL, d), OFDFT_EINVAL, "eval");
  expect_code(ofdft_eval_phi(P0, NULL, NULL, 1.0, 0.0, NULL, NULL, d), OFDFT_EINVAL, "eval_phi");
  expect_code(ofdft_hvp(P0, NULL, NULL, NULL, NULL, 0.0, NULL, d), OFDFT_EINVAL, "hvp");
  expect_code(ofdft_hvp_prepare(P0, NULL, NULL, NULL, 0.0, NULL, NULL, d), OFDFT_EINVAL, "hvp_prepare");
  expect_code(ofdft_potential_jvp(P0, NULL, NULL, NULL, NULL, NULL), OFDFT_EINVAL, "potential_jvp");
  expect_code(ofdft_hvp_fast(P0, NULL, NULL, NULL, NULL, d), OFDFT_EINVAL, "hvp_fast");
  expect_code(ofdft_tn_inner(P0, NULL, NULL, NULL, NULL, NULL, NULL, 1.0, d), OFDFT_EINVAL, "tn_inner");
  expect_code(ofdft_tn_direction(P0, NULL, NULL, NULL, NULL, NULL, 50, 4, iv, iv + 1, d), OFDFT_EINVAL,
              "tn_direction");
  expect_code(ofdft_fft_r2c(P0, NULL, NULL), OFDFT_EINVAL, "fft_r2c");
  expect_code(ofdft_fft_c2r(P0, NULL, NULL), OFDFT_EINVAL, "fft_c2r");
  expect_code(ofdft_convolve(P0, OFDFT_CONV_HARTREE, NULL, NULL), OFDFT_EINVAL, "convolve");
  expect_code(ofdft_vloc_build(P0, 0, NULL, 0, NULL, NULL, 1, 10, NULL), OFDFT_EINVAL, "vloc_build");
  expect_code(ofdft_local_forces(P0, NULL, d), OFDFT_EINVAL, "local_forces");
  expect_code(ofdft_ewald_forces(P0, 1, d, d, 1.0, 10.0, 0, 10, d), OFDFT_EINVAL, "ewald_forces");
  expect_code(ofdft_ewald(P0, 1, d, d, 1.0, 10.0, 0, 10, d, d + 1), OFDFT_EINVAL, "ewald");
  expect_code(ofdft_ewald_noalloc(P0, 1, d, d, 1.0, 10.0, 0, 10, d, d + 1), OFDFT_EINVAL, "ewald_noalloc");
  expect_code(ofdft_functional_stress(P0, NULL, 1.0, d), OFDFT_EINVAL, "functional_stress");
  expect_code(ofdft_local_stress(P0, NULL, d), OFDFT_EINVAL, "local_stress");
  expect_code(ofdft_ewald_stress(P0, 1, d, d, 1.0, 10.0, 0, 10, d, d + 6), OFDFT_EINVAL, "ewald_stress");
  expect_code(ofdft_gga(P0, 0, d, 0, NULL, NULL, NULL, 1.0, 0.0, NULL, d), OFDFT_EINVAL, "gga");
  expect_code(ofdft_lwt_set_kernel(P0, NULL, NULL), OFDFT_EINVAL, "lwt_set_kernel");
  expect_code(ofdft_lwt_range(P0, NULL, NULL, NULL, 1.0, 0.0, d), OFDFT_EINVAL, "lwt_range");
  expect_code(ofdft_lwt(P0, NULL, d, 0, NULL, NULL, NULL, 1.0, 0.0, NULL, d), OFDFT_EINVAL, "lwt");
  expect_code(ofdft_dot3(P0, NULL, NULL, d), OFDFT_EINVAL, "dot3");
  expect_code(ofdft_axpby(P0, 1.0, NULL, 1.0, NULL), OFDFT_EINVAL, "axpby");
  expect_code(ofdft_lincomb(P0, 1.0, NULL, 1.0, NULL, NULL), OFDFT_EINVAL, "lincomb");
  expect_code(ofdft_residual(P0, NULL, NULL, 0.0, NULL, d), OFDFT_EINVAL, "residual");
  expect_code(ofdft_cg_update(P0, 1.0, NULL, NULL, NULL, NULL, d), OFDFT_EINVAL, "cg_update");
  expect_code(ofdft_dot_vphip(P0, NULL, NULL, NULL, d), OFDFT_EINVAL, "dot_vphip");
  expect_code(ofdft_orth_sums(P0, NULL, NULL, NULL, d), OFDFT_EINVAL, "orth_sums");
  expect_code(ofdft_orth_normalize(P0, NULL, NULL, NULL), OFDFT_EINVAL, "orth_normalize");
  expect_code(ofdft_hvp_prepare_async(P0, NULL, NULL, NULL, 0.0, NULL, NULL), OFDFT_EINVAL, "hvp_prepare_async");
  expect_code(ofdft_plan_deferred_sums(P0, NULL), OFDFT_EINVAL, "plan_deferred_sums");
  expect_code(ofdft_convert(P0, NULL, OFDFT_F64, NULL, 8), OFDFT_EINVAL, "convert");

  /* batches */
  ofdft_batch B = NULL;
  expect_code(ofdft_batch_create(NULL, &P0, 1), OFDFT_EINVAL, "batch_create(out NULL)");
  expect_code(ofdft_batch_create(&B, NULL, 1), OFDFT_EINVAL, "batch_create(plans NULL)");
  expect_code(ofdft_batch_create(&B, &P0, 0), OFDFT_EINVAL, "batch_create(nb 0)");
  expect_code(ofdft_batch_create(&B, &P0, 1), OFDFT_EINVAL, "batch_create(null plan)");
  expect_code(ofdft_batch_destroy(NULL), 0, "batch_destroy(NULL) is a no-op");
  expect_code(ofdft_batch_tn_direction(NULL, vp, vp, vp, wp, wp, 50, iv, iv, d), OFDFT_EINVAL,
              "batch_tn_direction(NULL)");

  /* plan creation: argument errors */
  ofdft_plan P = NULL;
  expect_code(ofdft_plan_create(NULL, 32, 32, 32, lat, OFDFT_F64, 0, NULL), OFDFT_EINVAL, "create(out NULL)");
  expect_code(ofdft_plan_create(&P, 32, 32, 32, NULL, OFDFT_F64, 0, NULL), OFDFT_EINVAL, "create(lattice NULL)");
  expect_code(ofdft_plan_create(&P, 1, 32, 32, lat, OFDFT_F64, 0, NULL), OFDFT_EINVAL, "create(n0 = 1)");
  expect_code(ofdft_plan_create(&P, 32, 5000, 32, lat, OFDFT_F64, 0, NULL), OFDFT_EINVAL, "create(n1 = 5000)");
  expect_code(ofdft_plan_create(&P, 32, 32, 32, lat, 7, 0, NULL), OFDFT_EINVAL, "create(precision 7)");
  expect_code(ofdft_plan_create(&P, 34, 32, 32, lat, OFDFT_F64, 0, NULL), OFDFT_EINVAL, "create(n0 = 34 = 2 17)");
  expect_msg("products of 2, 3, 5, 7, 11 and 13", "factorisation message");
  expect_code(ofdft_plan_create(&P, 640, 32, 32, lat, OFDFT_F64, 0, NULL), OFDFT_EINVAL, "create(FP64 n0 = 640)");
  expect_msg("n0, n1 <= 601", "LDS bound message");
  expect_code(ofdft_plan_create(&P, 32, 32, 32, sing, OFDFT_F64, 0, NULL), OFDFT_EINVAL, "create(singular)");
  expect_msg("singular", "singular-lattice message");

  /* plan creation sweep: every length from 2 to 420 on each axis position,
   * both precisions, half and full grids; a plan is either created (with a
   * GPU) and destroyed at once, or refused with a message */
  int made = 0, refused = 0;
  for (int prec = 0; prec < 2; ++prec) {
    for (int full = 0; full < 2; ++full) {
      for (int n = 2; n <= 420; ++n) {
        const int dims[3][3] = {{n, 8, 8}, {8, n, 8}, {8, 8, n}};
        for (int a = 0; a < 3; ++a) {
          P = NULL;
          const int rc = ofdft_plan_create(&P, dims[a][0], dims[a][1], dims[a][2], lat,
                                           prec ? OFDFT_F32 : OFDFT_F64, full, NULL);
          ++nchecks;
          if (rc == 0) {
            if (!P || !smooth13(n)) {
              ++nfail;
              fprintf(stderr, "FAIL sweep n=%d: created a plan for a length it must refuse\n", n);
            }
            ++made;
            ofdft_plan_destroy(P);
          } else {
            ++refused;
            const char* msg = ofdft_last_error();
            if (P != NULL || !msg || !msg[0]) {
              ++nfail;
              fprintf(stderr, "FAIL sweep n=%d: rc=%d with plan %p, msg '%s'\n", n, rc, (void*)P, msg);
            }
            if (!smooth13(n) && !strstr(msg, "products of")) {
              ++nfail;
              fprintf(stderr, "FAIL sweep n=%d: refused for '%s', not the factorisation\n", n, msg);
            }
          }
        }
      }
    }
  }
  printf("sweep: %d plans created, %d refused\n", made, refused);
  if (nfail) {
    printf("%d of %d checks failed\n", nfail, nchecks);
    return 1;
  }
  printf("ok %d checks\n", nchecks);
  return 0;
}
