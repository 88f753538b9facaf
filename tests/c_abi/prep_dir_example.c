/*
 * prep_dir_example.c -- the reference's file-based flow through libmpkx,
 * from a plain C caller.
 *
 *   mpk2/driver.sh             -> the prep directory ROOT/cube16_8_6_1 with
 *                                 loop.g0, g0.part.8, l0, s1 and the values
 *                                 in ROOT/cube16.loop.val (written here with
 *                                 the ABI's generators and writers)
 *   mpi2_mpkwrtbufs DIR        -> mpkx_wrtbufs(DIR, 1): every rank%d.bufs
 *   mpi2_mpkexecbufs_val DIR   -> mpkx_execbufs_dir(DIR, NULL, vv) on the GPU
 *                                 (with --gpu): results.txt / results.bin
 *   verify_val                 -> k x SpMV of ones in plain C
 *                                 (mpk2/verify_val.c:44-75), compared with
 *                                 results.bin bitwise
 *
 * usage: prep_dir_example ROOT [--gpu]   (ROOT without '_': the reference
 * cuts the values path at the first '_' of the whole path)
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
    fprintf(stderr, "usage: %s ROOT [--gpu]\n", argv[0]);
    return 2;
  }
  const int gpu = argc > 2 && strcmp(argv[2], "--gpu") == 0;
  const int N = 16, k = 6, P = 8;
  char dir[4096], path[4200];
  snprintf(dir, sizeof dir, "%s/cube%d_%d_%d_1", argv[1], N, P, k);
  if (strchr(argv[1], '_')) {
    fprintf(stderr, "ROOT must not contain '_'\n");
    return 2;
  }

  /* driver.sh: graph, values, partition, levels and skirt files */
  int n;
  long nnz;
  CHECK(mpkx_gen2(4, N, 1.0, &n, &nnz, NULL, NULL, NULL));
  int *ptr = malloc(sizeof(int) * (n + 1)), *col = malloc(sizeof(int) * nnz);
  double *val = malloc(sizeof(double) * nnz);
  CHECK(mpkx_gen2(4, N, 1.0, &n, &nnz, ptr, col, val));
  snprintf(path, sizeof path, "mkdir -p %s", dir);
  if (system(path) != 0) return 1;
  snprintf(path, sizeof path, "%s/loop.g0", dir);
  CHECK(mpkx_write_crs(path, n, ptr, col));
  CHECK(mpkx_values_path(dir, path, sizeof path));
  CHECK(mpkx_write_vals(path, nnz, val));
  int dims[3] = {N, N, N}, blocks[3] = {2, 2, 2}, shift[3] = {0, 0, 0};
  int *part = malloc(sizeof(int) * n), *level = malloc(sizeof(int) * n);
  CHECK(mpkx_partition_grid(3, dims, blocks, shift, part));
  CHECK(mpkx_levels(n, ptr, col, part, NULL, level));
  long *sptr;
  int *srow, *sdist;
  CHECK(mpkx_comp_skirt(n, ptr, col, part, P, k, level, &sptr, &srow, &sdist));
  snprintf(path, sizeof path, "%s/g0.part.%d", dir, P);
  CHECK(mpkx_write_ints(path, n, part));
  snprintf(path, sizeof path, "%s/l0", dir);
  CHECK(mpkx_write_ints(path, n, level));
  snprintf(path, sizeof path, "%s/s1", dir);
  CHECK(mpkx_write_skirt(path, n, P, k, sptr, srow, sdist));

  /* mpi2_mpkwrtbufs on every rank */
  CHECK(mpkx_wrtbufs(dir, 1));
  printf("prep ok: %s/rank0..%d.bufs\n", dir, P - 1);
  if (!gpu) return 0;

  /* mpi2_mpkexecbufs_val on every rank (input: ones, as the reference) */
  double *vv = malloc(sizeof(double) * n * (k + 1));
  CHECK(mpkx_execbufs_dir(dir, NULL, vv));

  /* verify_val: k x SpMV of ones, CSR order, product and sum rounded apart */
  double *gold = malloc(sizeof(double) * n * (k + 1));
  for (int i = 0; i < n; i++) gold[i] = 1.0;
  for (int l = 0; l < k; l++)
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = ptr[i]; j < ptr[i + 1]; j++) {
        volatile double p = val[j] * gold[(long)l * n + col[j]];
        s += p;
      }
      gold[(long)(l + 1) * n + i] = s;
    }
  snprintf(path, sizeof path, "%s/results.bin", dir);
  FILE *f = fopen(path, "rb");
  if (!f) return 1;
  double *res = malloc(sizeof(double) * n * (k + 1));
  size_t got = fread(res, sizeof(double), (size_t)n * (k + 1), f);
  fclose(f);
  if (got != (size_t)n * (k + 1) || memcmp(res, gold, sizeof(double) * n * (k + 1)) != 0 ||
      memcmp(vv, gold, sizeof(double) * n * (k + 1)) != 0) {
    fprintf(stderr, "results.bin differs from k x SpMV\n");
    return 1;
  }
  printf("exec ok: results.bin bitwise equal to k x SpMV of ones (%d rows, k=%d)\n", n, k);
  return 0;
}
