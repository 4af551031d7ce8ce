/* A C caller of the boundary, compiled against include/pgfem3d_amd.h alone: it builds a
 * J2 box of quadratic tets, numbers its DOFs, builds the CSR pattern, creates the device
 * handle and runs one fused residual + tangent pass, then writes
 *   int64 ndof, int64 nnz, int64 status, double f[ndof], double vals[nnz]
 * to the file named by argv[1].  The mesh, supports, material and displacement are those of
 * tests/problems.py BoxProblem(2, PGF_TET10, j2_linear(70e3, 0.33, 250, 1000)) with
 * displacement(0.012, 0, seed) (no noise), so a test can check the file against the oracle.
 * Exit codes: 0 written, 3 no usable GPU (pgf_create gave PGF_ENODEV), 1 any other error. */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>

#include "pgfem3d_amd.h"

#define CHECK(call)                                                     \
  do {                                                                  \
    int rc_ = (call);                                                   \
    if (rc_ < 0) {                                                      \
      fprintf(stderr, "%s failed: %d\n", #call, rc_);                   \
      return 1;                                                         \
    }                                                                   \
  } while (0)

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 1;
  }
  const int32_t n = 2;
  const double stretch = 0.012;
  int64_t nn = 0, ne = 0, clen = 0;
  CHECK(pgf_mesh_box(n, PGF_TET10, 0.1, 1, &nn, &ne, &clen, NULL, NULL, NULL));
  double *X = malloc(sizeof(double) * 3 * nn);
  int64_t *eptr = malloc(sizeof(int64_t) * (ne + 1));
  int64_t *econn = malloc(sizeof(int64_t) * clen);
  CHECK(pgf_mesh_box(n, PGF_TET10, 0.1, 1, &nn, &ne, &clen, X, eptr, econn));
  pgf_mesh mesh = {nn, X, ne, eptr, econn, NULL, 0, NULL, NULL};

  /* bottom face fixed, top face u_z prescribed by value 1 */
  int64_t *id = calloc(3 * nn, sizeof(int64_t));
  for (int64_t k = 0; k < nn; ++k) {
    const double z = X[3 * k + 2];
    if (z < 1e-12) id[3 * k] = id[3 * k + 1] = id[3 * k + 2] = 1;
    if (z > 1.0 - 1e-12) id[3 * k + 2] = -1;
  }
  int64_t ndof = 0, nnz = 0;
  CHECK(pgf_generate_dof_ids(&mesh, id, &ndof));
  int32_t *Ap = malloc(sizeof(int32_t) * (ndof + 1));
  CHECK(pgf_psparse_apai(&mesh, id, ndof, Ap, NULL, &nnz));
  int32_t *Ai = malloc(sizeof(int32_t) * nnz);
  CHECK(pgf_psparse_apai(&mesh, id, ndof, Ap, Ai, &nnz));

  pgf_material mat = {0};
  mat.model = PGF_MODEL_J2_PLASTICITY;
  mat.devPotFlag = 1;
  mat.volPotFlag = 2;
  mat.E = 70e3;
  mat.nu = 0.33;
  mat.G = mat.E / (2.0 * (1.0 + mat.nu));
  mat.m10 = 0.5 * mat.G;
  mat.k0 = 250.0;
  mat.hp = 1000.0;

  pgf_handle h = NULL;
  const int rc = pgf_create(&mesh, id, ndof, Ap, Ai, 1, &mat, &h);
  if (rc == PGF_ENODEV) {
    printf("PGF_ENODEV\n");
    return 3;
  }
  CHECK(rc);

  /* u_z = s z, u_x = -0.3 s (x - 0.5) z, u_y = -0.3 s (y - 0.5) z on the free DOFs */
  double *r = calloc(ndof, sizeof(double));
  for (int64_t k = 0; k < nn; ++k) {
    const double *x = X + 3 * k;
    const double u[3] = {-0.3 * stretch * (x[0] - 0.5) * x[2], -0.3 * stretch * (x[1] - 0.5) * x[2],
                         stretch * x[2]};
    for (int c = 0; c < 3; ++c)
      if (id[3 * k + c] > 0) r[id[3 * k + c] - 1] = u[c];
  }
  const double defl = stretch;
  double *d_r = NULL, *d_defl = NULL, *d_f = NULL, *d_vals = NULL;
  if (hipMalloc((void **)&d_r, sizeof(double) * ndof) != hipSuccess ||
      hipMalloc((void **)&d_defl, sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&d_f, sizeof(double) * ndof) != hipSuccess ||
      hipMalloc((void **)&d_vals, sizeof(double) * nnz) != hipSuccess)
    return 1;
  if (hipMemcpy(d_r, r, sizeof(double) * ndof, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_defl, &defl, sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return 1;
  pgf_nodal_input in = {d_r, NULL, d_defl, NULL};
  const int64_t status = pgf_residual_tangent(h, &in, 1.0, d_f, d_vals);
  if (status < 0) {
    fprintf(stderr, "pgf_residual_tangent failed: %lld\n", (long long)status);
    return 1;
  }
  double *f = malloc(sizeof(double) * ndof), *vals = malloc(sizeof(double) * nnz);
  if (hipMemcpy(f, d_f, sizeof(double) * ndof, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(vals, d_vals, sizeof(double) * nnz, hipMemcpyDeviceToHost) != hipSuccess)
    return 1;
  CHECK(pgf_destroy(h));
  FILE *fp = fopen(argv[1], "wb");
  if (!fp) return 1;
  fwrite(&ndof, sizeof(int64_t), 1, fp);
  fwrite(&nnz, sizeof(int64_t), 1, fp);
  fwrite(&status, sizeof(int64_t), 1, fp);
  fwrite(f, sizeof(double), ndof, fp);
  fwrite(vals, sizeof(double), nnz, fp);
  fclose(fp);
  printf("ndof %lld nnz %lld status %lld\n", (long long)ndof, (long long)nnz, (long long)status);
  hipFree(d_r);
  hipFree(d_defl);
  hipFree(d_f);
  hipFree(d_vals);
  return 0;
}
