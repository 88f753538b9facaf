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
// every rank.  The shared-memory programs use the same directory with the
// graph in g0 and the values in <path up to the first '_'>.val
// (read_mpk, mpk2/mpkread.c:49-185; alloc_read_val, mpk2/mpkrun.c:12-32).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

using namespace mpkx;

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

}  // extern "C"

namespace mpkx {

PrepDir::~PrepDir() {
  mpkx_free(ptr);
  mpkx_free(col);
  mpkx_free(sptr);
  mpkx_free(srow);
  mpkx_free(sdist);
}

int load_prep_dir(const char *dir, const char *graph_file, const char *values_suffix, bool with_values,
                  PrepDir &pd) {
  int rc;
  if ((rc = mpkx_dir_params(dir, &pd.npart, &pd.nlevel, &pd.nphase))) return rc;
  const std::string d(dir);
  if ((rc = mpkx_read_crs((d + "/" + graph_file).c_str(), &pd.n, &pd.ptr, &pd.col))) return rc;
  const int n = pd.n, nph = pd.nphase > 0 ? pd.nphase : 1;
  pd.parts.assign(nph, std::vector<int>(n));
  pd.levels.assign(pd.nphase, std::vector<int>(n));
  for (int ph = 0; ph < nph; ph++) {
    const std::string f = d + "/g" + std::to_string(ph) + ".part." + std::to_string(pd.npart);
    if ((rc = mpkx_read_ints(f.c_str(), n, pd.parts[ph].data(), nullptr))) return rc;
  }
  for (int ph = 0; ph < pd.nphase; ph++) {
    const std::string f = d + "/l" + std::to_string(ph);
    if ((rc = mpkx_read_ints(f.c_str(), n, pd.levels[ph].data(), nullptr))) return rc;
  }
  if ((rc = mpkx_read_skirt((d + "/s" + std::to_string(pd.nphase)).c_str(), n, pd.npart, pd.nlevel, &pd.sptr,
                            &pd.srow, &pd.sdist)))
    return rc;
  if (with_values) {
    // the path up to its first '_', then the suffix
    const char *u = strchr(dir, '_');
    if (!u) return fail(MPKX_EINVAL, "no '_' in %s", dir);
    const std::string vpath = std::string(dir, (size_t)(u - dir)) + values_suffix;
    pd.val.resize((size_t)pd.ptr[n]);
    if ((rc = mpkx_read_vals(vpath.c_str(), pd.ptr[n], pd.val.data()))) return rc;
  }
  return MPKX_OK;
}

}  // namespace mpkx

extern "C" {

int mpkx_wrtbufs(const char *dir, int with_values) {
  clear_error();
  PrepDir pd;
  int rc;
  if ((rc = load_prep_dir(dir, "loop.g0", ".loop.val", with_values != 0, pd))) return rc;
  const int nph = pd.nphase > 0 ? pd.nphase : 1;
  std::vector<int *> pp(nph);
  std::vector<const int *> lp(pd.nphase);
  for (int ph = 0; ph < nph; ph++) pp[ph] = pd.parts[ph].data();
  for (int ph = 0; ph < pd.nphase; ph++) lp[ph] = pd.levels[ph].data();
  mpkx_buffers **bufs = nullptr;
  if ((rc = mpkx_fill_buffers(pd.n, pd.ptr, pd.col, with_values ? pd.val.data() : nullptr, pd.nlevel, pd.nphase,
                              pp.data(), pd.nphase ? lp.data() : nullptr, pd.sptr, pd.srow, pd.sdist, pd.npart,
                              &bufs)))
    return rc;
  for (int r = 0; r < pd.npart && rc == MPKX_OK; r++) rc = mpkx_write_buffers(bufs[r], dir);
  mpkx_del_bufs_array(bufs, pd.npart);
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
