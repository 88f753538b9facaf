/*
 * ref_exec_harness.c -- builds the reference MPI exec program
 * (mpk2/mpi2_mpkexecbufs_val.c) unchanged, from where it lies in the
 * reference checkout, and adds a driver that loads a seeded input vector
 * instead of the reference's hard-coded ones (mpi2_mpkexecbufs_val.c:86-87).
 * Test infrastructure only (oracle).
 *
 * usage: mpirun -n P ref_exec DIR XFILE REPS
 *   XFILE: n doubles (native endian); REPS: timed repetitions (min kept,
 *   each bracketed by MPI_Barrier as in mpi2_mpkexecbufs_val.c:89-100).
 * Output: DIR/results.bin (n*(nlevel+1) doubles, level-major) and one line
 *   "ref_exec min_seconds=<t> ranks=<P>" on rank 0.
 */
#define main mpk2_reference_main
#include "mpi2_mpkexecbufs_val.c"
#undef main

int main(int argc, char *argv[]) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s DIR XFILE REPS\n", argv[0]);
    return 1;
  }
  MPI_Init(&argc, &argv);
  int rank, size;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &size);
  buffers_t *bufs = read_buffers(argv[1], rank);
  alloc_val_bufs(bufs);
  double *x = (double *)malloc(sizeof(double) * bufs->n);
  FILE *fx = fopen(argv[2], "rb");
  if (fx == NULL || fread(x, sizeof(double), bufs->n, fx) != (size_t)bufs->n) {
    fprintf(stderr, "cannot read %s\n", argv[2]);
    MPI_Abort(MPI_COMM_WORLD, 2);
  }
  fclose(fx);
  int reps = atoi(argv[3]);
  if (reps < 1) reps = 1;
  double best = 0.0;
  for (int t = 0; t < reps; t++) {
    /* re-load the input each repetition so every repetition does the same work */
    for (int i = 0; i < bufs->rcount[0]; i++) bufs->vv_buf[i] = x[bufs->idx_buf[i]];
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    mpk_exec_bufs_val(bufs);
    MPI_Barrier(MPI_COMM_WORLD);
    double dt = MPI_Wtime() - t0;
    if (t == 0 || dt < best) best = dt;
  }
  double *vv = (double *)malloc(sizeof(double) * (size_t)bufs->n * (bufs->nlevel + 1));
  collect_results(bufs, vv);
  if (rank == 0) {
    char fname[4096];
    snprintf(fname, sizeof fname, "%s/results.bin", argv[1]);
    FILE *f = fopen(fname, "wb");
    fwrite(vv, sizeof(double), (size_t)bufs->n * (bufs->nlevel + 1), f);
    fclose(f);
    printf("ref_exec min_seconds=%.9f ranks=%d\n", best, size);
  }
  free(vv);
  free(x);
  del_bufs(bufs);
  MPI_Finalize();
  return 0;
}
