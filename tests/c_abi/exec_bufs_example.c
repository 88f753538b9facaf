/*
 * exec_bufs_example.c -- a reference-style C caller of libmpkx.
 *
 * It does what mpk2/driver.sh + mpi2_mpkwrtbufs and then
 * `mpirun -n P mpi2_mpkexecbufs_val` do (mpk2/mpi2_mpkexecbufs_val.c:77-121),
 * through the C ABI only: generate the 3D 7-point matrix, partition it,
 * compute levels and the skirt, fill every rank's buffers, write and read
 * them back in the rank%d.bufs format, and with --gpu run all ranks'
 * phases on the GPU (mpkx_plan_create + mpkx_exec_bufs), collect like
 * collect_results and compare bitwise with k x SpMV computed here in
 * plain C (verify_val's loop, mpk2/verify_val.c:44-75).
 *
 * usage: exec_bufs_example DIR [--gpu]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpkx.h"

#define CHECK(call)                                                   \
  do {                                                                \
    if ((call) != MPKX_OK) {                                          \
      fprintf(stderr, "%s failed: %s\n", #call, mpkx_last_error());   \
      return 1;                                                       \
    }                                                                 \
  } while (0)

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s DIR [--gpu]\n", argv[0]);
    return 2;
  }
  const char *dir = argv[1];
  const int gpu = argc > 2 && strcmp(argv[2], "--gpu") == 0;
  const int N = 16, k = 6, nphase = 1, P = 8;
  int n;
  long nnz;
  CHECK(mpkx_gen2(4, N, -1.0, &n, &nnz, NULL, NULL, NULL));
  int *ptr = malloc(sizeof(int) * (n + 1)), *col = malloc(sizeof(int) * nnz);
  double *val = malloc(sizeof(double) * nnz);
  CHECK(mpkx_gen2(4, N, -1.0, &n, &nnz, ptr, col, val));

  /* partition (gP.part.N), levels (lP), skirt (sN) */
  int dims[3] = {N, N, N}, blocks[3] = {2, 2, 2}, shift[3] = {0, 0, 0};
  int *part = malloc(sizeof(int) * n), *level = malloc(sizeof(int) * n);
  CHECK(mpkx_partition_grid(3, dims, blocks, shift, part));
  CHECK(mpkx_levels(n, ptr, col, part, NULL, level));
  long *sptr;
  int *srow, *sdist;
  CHECK(mpkx_comp_skirt(n, ptr, col, part, P, k, level, &sptr, &srow, &sdist));

  /* mpi2_mpkwrtbufs: every rank's buffers, written and read back */
  int *parts[1] = {part};
  const int *levels[1] = {level};
  mpkx_buffers **bufs;
  CHECK(mpkx_fill_buffers(n, ptr, col, val, k, nphase, parts, levels, sptr, srow, sdist, P, &bufs));
  long total = 0;
  mpkx_buffers *back[8];
  for (int r = 0; r < P; r++) {
    CHECK(mpkx_write_buffers(bufs[r], dir));
    back[r] = mpkx_read_buffers(dir, r);
    if (!back[r]) {
      fprintf(stderr, "read_buffers(%d): %s\n", r, mpkx_last_error());
      return 1;
    }
    if (back[r]->buf_count != bufs[r]->buf_count || back[r]->mcol_count != bufs[r]->mcol_count ||
        memcmp(back[r]->mcol_buf, bufs[r]->mcol_buf, sizeof(long) * bufs[r]->mcol_count) != 0) {
      fprintf(stderr, "rank %d: buffers differ after the file round trip\n", r);
      return 1;
    }
    total += back[r]->buf_count;
  }
  printf("prep ok: %d ranks, %ld vv_buf slots, rank%%d.bufs round trip equal\n", P, total);
  if (!gpu) return 0;

  /* the exec: vv_buf[0 .. rcount[0]) = x[idx_buf[i]] as the reference main loads it */
  double *x = malloc(sizeof(double) * n);
  mpkx_uniform_pm1(0x5EED, n, x);
  for (int r = 0; r < P; r++) {
    CHECK(mpkx_alloc_val_bufs(back[r]));
    for (int i = 0; i < back[r]->rcount[0]; i++) back[r]->vv_buf[i] = x[back[r]->idx_buf[i]];
  }
  mpkx_plan *plan;
  CHECK(mpkx_plan_create(&plan, back, P));
  CHECK(mpkx_exec_bufs(plan, back));
  mpkx_plan_destroy(plan);
  /* collect_results: vv[idx_buf[i]] = vv_buf[i] */
  double *vv = calloc((size_t)n * (k + 1), sizeof(double));
  for (int r = 0; r < P; r++)
    for (int i = 0; i < back[r]->buf_count; i++) vv[back[r]->idx_buf[i]] = back[r]->vv_buf[i];
  /* gold: k x SpMV in CSR order, product and sum rounded separately */
  double *gold = malloc(sizeof(double) * (size_t)n * (k + 1));
  memcpy(gold, x, sizeof(double) * n);
  for (int l = 0; l < k; l++)
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = ptr[i]; j < ptr[i + 1]; j++) {
        double prod = val[j] * gold[(size_t)l * n + col[j]];
        s = s + prod;
      }
      gold[(size_t)(l + 1) * n + i] = s;
    }
  if (memcmp(vv, gold, sizeof(double) * (size_t)n * (k + 1)) != 0) {
    fprintf(stderr, "exec result differs from k x SpMV\n");
    return 1;
  }
  printf("exec ok: %d ranks on the GPU, Krylov block bitwise equal to k x SpMV (n=%d, k=%d)\n", P, n, k);
  for (int r = 0; r < P; r++) mpkx_del_bufs(back[r]);
  mpkx_del_bufs_array(bufs, P);
  mpkx_free(sptr), mpkx_free(srow), mpkx_free(sdist);
  free(ptr), free(col), free(val), free(part), free(level), free(x), free(vv), free(gold);
  return 0;
}
