/* A C99 caller of the C ABI (include/cles.h), compiled and linked against
 * libcles.so by tests/test_c_abi.py.  It uses only host-side entry points:
 * the reference defaults, the option validation cles_create runs before it
 * touches a device, and the PETSc binary Vec writer/reader.  It prints the
 * struct layout for the ctypes mirror to compare with. */
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "cles.h"

static int fails = 0;
#define CHECK(cond)                                             \
  do {                                                          \
    if (!(cond)) {                                              \
      printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);     \
      fails++;                                                  \
    }                                                           \
  } while (0)

int main(int argc, char **argv) {
  cles_params p;
  cles_default_params(&p, 8, 9, 10);
  /* init.c:85-96, rhs.c:2581-2585, les.c:151 */
  CHECK(p.mx == 8 && p.my == 9 && p.mz == 10);
  CHECK(p.ren == 3000. && p.Ma == 2.81 && p.Pr == 0.72 && p.gamma == 1.4 && p.dt == 0.01 && p.st == 1.);
  CHECK(p.fluxorder == 3 && p.eno_order == 2 && p.central == 2 && p.max_cs == 0.05);
  CHECK(p.precision == CLES_PREC_FP64 && p.calc_sigma == 0 && p.i_homo_filter == 0);
  for (int q = 0; q < 6; q++) CHECK(p.bctype[q] == 7);

  /* validation before any device work: a NULL context and the code */
  double coor[8 * 9 * 10 * 3];
  memset(coor, 0, sizeof(coor));
  int err = 0;
  cles_params bad = p;
  bad.fluxorder = 4;
  CHECK(cles_create(&bad, coor, &err) == NULL && err == CLES_EINVAL);
  CHECK(strstr(cles_last_error(NULL), "fluxorder") != NULL);
  bad = p;
  bad.bctype[1] = 3; /* periodic on one side only */
  CHECK(cles_create(&bad, coor, &err) == NULL && err == CLES_EINVAL);
  bad = p;
  bad.les = 2;
  bad.invicid = 1;
  CHECK(cles_create(&bad, coor, &err) == NULL && err == CLES_EINVAL);
  bad = p;
  bad.mx = 3;
  CHECK(cles_create(&bad, coor, &err) == NULL && err == CLES_EINVAL);

  /* PETSc binary Vec round trip (host only) */
  if (argc > 1) {
    double x[37], y[37];
    for (int n = 0; n < 37; n++) x[n] = (n - 18) * 0.375 + 1e-300 * n;
    CHECK(cles_petsc_vec_write(argv[1], x, 37, 5) == CLES_OK);
    CHECK(cles_petsc_vec_read(argv[1], y, 37) == CLES_OK);
    CHECK(memcmp(x, y, sizeof(x)) == 0);
    CHECK(cles_petsc_vec_read(argv[1], y, 36) == CLES_EINVAL);
  }
  CHECK(cles_footprint(NULL) == 0);

  /* layout for the ctypes mirror */
  printf("sizeof cles_params %zu\n", sizeof(cles_params));
  printf("offsetof ren %zu les %zu precision %zu averaging %zu k_homo_filter %zu\n",
         offsetof(cles_params, ren), offsetof(cles_params, les), offsetof(cles_params, precision),
         offsetof(cles_params, averaging), offsetof(cles_params, k_homo_filter));
  printf("sizeof cles_ibm_info %zu offsetof pmin %zu cr3 %zu\n", sizeof(cles_ibm_info),
         offsetof(cles_ibm_info, pmin), offsetof(cles_ibm_info, cr3));
  printf("sizeof cles_vtk_opts %zu offsetof theta %zu pres %zu airfoil %zu\n", sizeof(cles_vtk_opts),
         offsetof(cles_vtk_opts, theta), offsetof(cles_vtk_opts, pres), offsetof(cles_vtk_opts, airfoil));
  printf("%s %d\n", fails ? "FAILED" : "OK", fails);
  return fails != 0;
}
