// host_dirs.cpp -- the reference's prep and exec directories.
//
// The reference keeps the output of its prep in a directory named
// <head>_<npart>_<nlevel>_<nphase> (mpk2/comm_data.c:27-43):
//   loop.g0             the graph (read_matrix, comm_data.c:110-120)
//   g<ph>.part.<npart>  the partition of phase ph (fill_part, :60-75);
//                       with nphase = 0 only g0.part.<npart>
//   l<ph>               the levels of phase ph < nphase (fill_level, :77-91)
//   s<nphase>           the skirt (fill_skirt, :93-107)
// and the matrix values beside it, in <path up to the first '_'>.loop.val
// (alloc_read_val, :122-149).  mpi2_mpkwrtbufs reads that directory once
// per MPI rank and writes <dir>/rank<r>.bufs (mpk2/mpi2_mpkwrtbufs.c:43-63);
// mpi2_mpkexecbufs_val writes <dir>/results.txt and results.bin
// (mpk2/mpi2_mpkexecbufs_val.c:58-75).  Here one call does the prep of
// every rank.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

using namespace mpkx;

namespace {

// Frees what the readers allocate.
struct Owned {
  std::vector<void *> p;
  ~Owned() {
    for (void *q : p) mpkx_free(q);
  }
};

}  // namespace

extern "C" {

int mpkx_dir_params(const char *dir, int *npart, int *nlevel, int *nphase) {
  clear_error();
  if (!dir) return fail(MPKX_EINVAL, "dir_params: null directory");
  // the third '_' from the end starts "_npart_nlevel_nphase" (read_dir)
  long j = (long)strlen(dir), seen = 0;
  for (; j >= 0; j--) {
    if (dir[j] == '_') seen++;
    if (seen == 3) break;
  }
  int p = -1, l = -1, ph = -1;
  if (seen != 3 || j < 0 || sscanf(dir + j, "_%d_%d_%d", &p, &l, &ph) != 3 || p <= 0 || l <= 0 || ph < 0)
    return fail(MPKX_EINVAL, "unknown directory naming %s (expecting ghead_npart_nlevel_nphase)", dir);
  if (npart) *npart = p;
  if (nlevel) *nlevel = l;
  if (nphase) *nphase = ph;
  return MPKX_OK;
}

int mpkx_values_path(const char *dir, char *out, size_t size) {
  clear_error();
  if (!dir || !out) return fail(MPKX_EINVAL, "values_path: bad arguments");
  // alloc_read_val: the path up to its first '_', then ".loop.val"
  const char *u = strchr(dir, '_');
  if (!u) return fail(MPKX_EINVAL, "values_path: no '_' in %s", dir);
  const std::string path = std::string(dir, (size_t)(u - dir)) + ".loop.val";
  if (path.size() + 1 > size) return fail(MPKX_EINVAL, "values_path: buffer too small");
  memcpy(out, path.c_str(), path.size() + 1);
  return MPKX_OK;
}

int mpkx_wrtbufs(const char *dir, int with_values) {
  clear_error();
  int npart, nlevel, nphase, rc;
  if ((rc = mpkx_dir_params(dir, &npart, &nlevel, &nphase))) return rc;
  const std::string d(dir);
  Owned own;
  int n = 0, *ptr = nullptr, *col = nullptr;
  if ((rc = mpkx_read_crs((d + "/loop.g0").c_str(), &n, &ptr, &col))) return rc;
  own.p.push_back(ptr);
  own.p.push_back(col);
  const int nph = nphase > 0 ? nphase : 1;
  std::vector<std::vector<int>> parts(nph, std::vector<int>(n)), levels(nphase, std::vector<int>(n));
  for (int ph = 0; ph < nph; ph++) {
    const std::string f = d + "/g" + std::to_string(ph) + ".part." + std::to_string(npart);
    if ((rc = mpkx_read_ints(f.c_str(), n, parts[ph].data(), nullptr))) return rc;
  }
  for (int ph = 0; ph < nphase; ph++) {
    const std::string f = d + "/l" + std::to_string(ph);
    if ((rc = mpkx_read_ints(f.c_str(), n, levels[ph].data(), nullptr))) return rc;
  }
  long *sptr = nullptr;
  int *srow = nullptr, *sdist = nullptr;
  if ((rc = mpkx_read_skirt((d + "/s" + std::to_string(nphase)).c_str(), n, npart, nlevel, &sptr, &srow, &sdist)))
    return rc;
  own.p.push_back(sptr);
  own.p.push_back(srow);
  own.p.push_back(sdist);
  std::vector<double> val;
  if (with_values) {
    char vpath[4096];
    if ((rc = mpkx_values_path(dir, vpath, sizeof vpath))) return rc;
    val.resize((size_t)ptr[n]);
    if ((rc = mpkx_read_vals(vpath, ptr[n], val.data()))) return rc;
  }
  std::vector<int *> pp(nph);
  std::vector<const int *> lp(nphase);
  for (int ph = 0; ph < nph; ph++) pp[ph] = parts[ph].data();
  for (int ph = 0; ph < nphase; ph++) lp[ph] = levels[ph].data();
  mpkx_buffers **bufs = nullptr;
  if ((rc = mpkx_fill_buffers(n, ptr, col, with_values ? val.data() : nullptr, nlevel, nphase, pp.data(),
                              nphase ? lp.data() : nullptr, sptr, srow, sdist, npart, &bufs)))
    return rc;
  for (int r = 0; r < npart && rc == MPKX_OK; r++) rc = mpkx_write_buffers(bufs[r], dir);
  mpkx_del_bufs_array(bufs, npart);
  return rc;
}

int mpkx_write_results(const char *dir, int n, int nlevel, const double *vv) {
  clear_error();
  if (!dir || n <= 0 || nlevel <= 0 || !vv) return fail(MPKX_EINVAL, "write_results: bad arguments");
  const std::string d(dir);
  FILE *f = fopen((d + "/results.txt").c_str(), "w");
  if (!f) return fail(MPKX_EIO, "cannot open %s/results.txt", dir);
  for (int l = 0; l < nlevel; l++) {  // levels 0 .. nlevel-1, as the reference writes them
    for (int i = 0; i < n; i++) fprintf(f, "%lf ", vv[(long)n * l + i]);
    fprintf(f, "\n");
  }
  fclose(f);
  f = fopen((d + "/results.bin").c_str(), "wb");
  if (!f) return fail(MPKX_EIO, "cannot open %s/results.bin", dir);
  const size_t count = (size_t)n * (nlevel + 1);
  const bool ok = fwrite(vv, sizeof(double), count, f) == count;
  fclose(f);
  return ok ? MPKX_OK : fail(MPKX_EIO, "write_results: short write");
}

}  // extern "C"
