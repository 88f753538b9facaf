// host_buffers.cpp -- exec-program construction for every partition at once.
//
// Produces, for every rank, the same rank%d.bufs bytes as the reference
// prep (mpk2/buffers.c:573-586 + mpk2/mpi2_mpkwrtbufs.c:12-30).  The
// reference runs the construction once per MPI rank over a dense table
// char[nlevel*n*npart*npart] (mpk2/buffers.c:21-29), which overflows `int`
// beyond about a dozen partitions at 128^3.  Here the communication table is
// a sorted set of (source, target, vv index) keys built once per phase for
// all ranks together; every rank's counts, displacements and index lists
// are then read off that set.  The two construction passes, the store_part
// bookkeeping, the LAPJV relabelling of phases >= 1 and the level-major
// iteration order are kept exactly, so the output is byte-identical.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "internal.h"

extern "C" mpkx_buffers *mpkx_new_bufs_internal(int n, int npart, int nlevel, int nphase, int rank);

using namespace mpkx;

namespace {

struct Ctx {
  int n, np, nlevel, nphase;
  const int *ptr, *col;
  int **part;  // part[ph], relabelled in place
  const int *const *lev;
  SkirtView sk;
  std::vector<int> sp;  // store_part[n*(nlevel+1)]
  int err = 0;
};

inline uint64_t key(int s, int t, int np, long v) {
  return ((uint64_t)((long)s * np + t) << 32) | (uint64_t)v;
}

int prev_level(const Ctx &c, int ph, int i) { return ph ? c.lev[ph - 1][i] : 0; }

// min_or_0 (mpk2/buffers.c:68-77)
int lmin_prev(const Ctx &c, int phm1) {
  if (c.nphase == 0 || phm1 == -1) return 0;
  const int *l = c.lev[phm1];
  int m = l[0];
  for (int i = 0; i < c.n; i++) m = std::min(m, l[i]);
  return m;
}

// max_or_nlevel (mpk2/buffers.c:79-90)
int lmax_cap(const Ctx &c, int ph) {
  if (ph == c.nphase) return c.nlevel;
  const int *l = c.lev[ph];
  int m = l[0];
  for (int i = 1; i < c.n; i++) m = std::max(m, l[i]);
  return std::min(m, c.nlevel);
}

// proc_vertex (mpk2/buffers.c:53-66): sources of (level, i) held by another
// partition than `tgt` are communicated to it.
void mark_sources(Ctx &c, std::vector<uint64_t> &keys, int tgt, int i, int level) {
  for (int j = c.ptr[i]; j < c.ptr[i + 1]; j++) {
    long sv = (long)c.n * (level - 1) + c.col[j];
    int s = c.sp[sv];
    if (s == -1) {
      c.err = 1;
      continue;
    }
    if (s != tgt) keys.push_back(key(s, tgt, c.np, sv));
  }
}

void add_init(const Ctx &c, std::vector<uint64_t> &keys) {  // buffers.c:45-51
  for (int i = 0; i < c.n; i++) keys.push_back(key(c.part[0][i], c.part[0][i], c.np, i));
}

// phase_comm_table (mpk2/buffers.c:125-153); first=true is the sizing pass
// that only takes unassigned values and records who computes them.
void phase_table(Ctx &c, int ph, bool first, std::vector<uint64_t> &keys) {
  const int *ll = c.lev[ph];
  int hi = lmax_cap(c, ph);
  for (int level = lmin_prev(c, ph - 1) + 1; level <= hi; level++)
    for (int i = 0; i < c.n; i++) {
      long iv = (long)c.n * level + i;
      if (prev_level(c, ph, i) < level && level <= ll[i] && (!first || c.sp[iv] == -1)) {
        int cp = c.part[ph][i];
        mark_sources(c, keys, cp, i, level);
        if (first) c.sp[iv] = cp;
      }
    }
}

// skirt_comm_table (mpk2/buffers.c:155-179), partitions in order.
void skirt_table(Ctx &c, std::vector<uint64_t> &keys) {
  int lo = lmin_prev(c, c.nphase - 1) + 1;
  for (int p = 0; p < c.np; p++)
    for (int level = lo; level <= c.nlevel; level++)
      for (long t = c.sk.sptr[p]; t < c.sk.sptr[p + 1]; t++) {
        int i = c.sk.srow[t], d = c.sk.sdist[t];
        if (prev_level(c, c.nphase, i) < level && d >= 0 && level <= c.nlevel - d) {
          mark_sources(c, keys, p, i, level);
          c.sp[(long)c.n * level + i] = p;
        }
      }
}

// Sorted, duplicates removed.  LSD radix sort on 11-bit digits over the
// bits in use (the pair s*np+t above bit 32, the vv index below): at 128^3
// with 64 partitions a table holds about 18M keys, which std::sort took
// most of the prep time to order.
void sort_unique(std::vector<uint64_t> &k) {
  uint64_t used = 0;
  for (uint64_t x : k) used |= x;
  constexpr int D = 11;
  constexpr size_t NB = (size_t)1 << D;
  std::vector<uint64_t> tmp(k.size());
  std::vector<size_t> cnt(NB);
  for (int shift = 0; shift < 64 && (used >> shift) != 0; shift += D) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (uint64_t x : k) cnt[(x >> shift) & (NB - 1)]++;
    size_t sum = 0;
    for (size_t b = 0; b < NB; b++) {
      const size_t c = cnt[b];
      cnt[b] = sum;
      sum += c;
    }
    for (uint64_t x : k) tmp[cnt[(x >> shift) & (NB - 1)]++] = x;
    k.swap(tmp);
  }
  k.erase(std::unique(k.begin(), k.end()), k.end());
}

// reduce_comm (mpk2/reduce_comm.c:31-82): collect_comm's table, then the
// LAPJV permutation relabels the new phase's partitions.
int relabel(Ctx &c, int ph) {
  std::vector<uint64_t> keys;
  const int *ll = c.lev[ph];
  int hi = lmax_cap(c, ph);
  for (int level = lmin_prev(c, ph - 1) + 1; level <= hi; level++)
    for (int i = 0; i < c.n; i++) {
      if (!(prev_level(c, ph, i) < level && level <= ll[i])) continue;
      int tgt = c.part[ph][i];
      for (int j = c.ptr[i]; j < c.ptr[i + 1]; j++) {
        long sv = (long)c.n * (level - 1) + c.col[j];
        int s = c.sp[sv];
        if (s != -1) keys.push_back(key(s, tgt, c.np, sv));
      }
    }
  sort_unique(keys);
  std::vector<int> sums((size_t)c.np * c.np, 0), perm(c.np);
  for (uint64_t k : keys) sums[k >> 32]++;
  int rc = mpkx_lapjv(sums.data(), c.np, perm.data());
  if (rc) return rc;
  for (int i = 0; i < c.n; i++) c.part[ph][i] = perm[c.part[ph][i]];
  return MPKX_OK;
}

// Per rank, the communicated vv indices in the order fill_idx_rsbuf writes
// them (mpk2/buffers.c:354-377): receives by (source, index), sends by
// (target, index).
struct PhaseComm {
  std::vector<std::vector<int>> recv, send;  // [rank] -> vv indices
  std::vector<int> cnt;                      // [s*np+t]
};

void split_comm(const Ctx &c, const std::vector<uint64_t> &keys, PhaseComm &pc) {
  pc.recv.assign(c.np, {});
  pc.send.assign(c.np, {});
  pc.cnt.assign((size_t)c.np * c.np, 0);
  for (uint64_t k : keys) {
    long st = (long)(k >> 32);
    int s = (int)(st / c.np), t = (int)(st % c.np);
    int v = (int)(k & 0xffffffffu);
    pc.cnt[st]++;
    pc.recv[t].push_back(v);
    pc.send[s].push_back(v);
  }
}

// iterator (mpk2/buffers.c:218-238) with phase_cond / skirt_cond: the vv
// indices level*n+i every rank computes in phase ph, level-major.  The
// reference's get_prevlmin always yields 0 (a shadowed local), so levels
// start at 1.
void mbuf_lists(const Ctx &c, int ph, std::vector<std::vector<long>> &mb) {
  mb.assign(c.np, {});
  int hi = lmax_cap(c, ph);
  if (ph < c.nphase) {
    for (int level = 1; level <= hi; level++)
      for (int i = 0; i < c.n; i++)
        if (prev_level(c, ph, i) < level && level <= c.lev[ph][i])
          mb[c.part[ph][i]].push_back((long)level * c.n + i);
  } else {
    for (int p = 0; p < c.np; p++)
      for (int level = 1; level <= hi; level++)
        for (long t = c.sk.sptr[p]; t < c.sk.sptr[p + 1]; t++) {
          int i = c.sk.srow[t], d = c.sk.sdist[t];
          if (prev_level(c, c.nphase, i) < level && 0 <= d && level <= c.nlevel - d)
            mb[p].push_back((long)level * c.n + i);
        }
  }
}

}  // namespace

extern "C" {

int mpkx_fill_buffers(int n, const int *ptr, const int *col, const double *val, int nlevel,
                      int nphase, int **parts, const int *const *levels, const long *sptr,
                      const int *srow, const int *sdist, int npart, mpkx_buffers ***out) {
  clear_error();
  if (n <= 0 || nlevel <= 0 || nphase < 0 || npart <= 0 || !parts || !sptr || !out)
    return fail(MPKX_EINVAL, "fill_buffers: bad arguments");
  if ((long)n * (nlevel + 1) > INT_MAX) return fail(MPKX_EINVAL, "fill_buffers: n*(nlevel+1) exceeds int");
  if (nphase > 0 && !levels) return fail(MPKX_EINVAL, "fill_buffers: levels missing");
  for (int ph = 0; ph < std::max(nphase, 1); ph++)
    for (int i = 0; i < n; i++)
      if (parts[ph][i] < 0 || parts[ph][i] >= npart)
        return fail(MPKX_EINVAL, "fill_buffers: phase %d label %d out of range", ph, parts[ph][i]);
  Ctx c;
  c.n = n;
  c.np = npart;
  c.nlevel = nlevel;
  c.nphase = nphase;
  c.ptr = ptr;
  c.col = col;
  c.part = parts;
  c.lev = levels;
  c.sk = {sptr, srow, sdist};
  long spn = (long)n * (nlevel + 1);
  MPKX_TRY_ALLOC(c.sp.assign(spn, -1));
  for (int i = 0; i < n; i++) c.sp[i] = parts[0][i];

  const int nph1 = nphase + 1;
  std::vector<std::vector<int>> mcount(nph1, std::vector<int>(npart, 0));
  std::vector<std::vector<int>> cnt1(nph1);  // pass-1 counts [s*np+t]

  // ---- pass 1: sizes (fill_bufsize_rscount_displs, buffers.c:284-315)
  try {
    for (int ph = 0; ph < nphase; ph++) {
      if (ph > 0) {
        int rc = relabel(c, ph);
        if (rc) return rc;
      }
      std::vector<uint64_t> keys;
      if (ph == 0) add_init(c, keys);
      phase_table(c, ph, true, keys);
      sort_unique(keys);
      cnt1[ph].assign((size_t)npart * npart, 0);
      for (uint64_t k : keys) cnt1[ph][k >> 32]++;
    }
    // The final (skirt) table is built once, in pass 2: the phase tables
    // never read values computed in the final phase (a schedule that does
    // fails with c.err), and the skirt table sets the same store_part
    // entries in the same order whether or not it ran before, so a sizing
    // run of it would only repeat pass 2's keys.
  } catch (const std::bad_alloc &) {
    return fail(MPKX_ENOMEM, "fill_buffers: pass 1 allocation");
  }
  if (c.err) return fail(MPKX_ESTRUCT, "fill_buffers: a source value is never stored (schedule inconsistent)");

  // ---- pass 2: contents (fill_bufs, buffers.c:392-411)
  std::vector<PhaseComm> comm(nph1);
  std::vector<std::vector<std::vector<long>>> mb(nph1);
  try {
    for (int ph = 0; ph < nphase; ph++) {
      std::vector<uint64_t> keys;
      if (ph == 0) add_init(c, keys);
      phase_table(c, ph, false, keys);
      sort_unique(keys);
      split_comm(c, keys, comm[ph]);
      mbuf_lists(c, ph, mb[ph]);
    }
    std::vector<uint64_t> keys;
    if (nphase == 0) add_init(c, keys);
    skirt_table(c, keys);
    sort_unique(keys);
    split_comm(c, keys, comm[nphase]);
    mbuf_lists(c, nphase, mb[nphase]);
  } catch (const std::bad_alloc &) {
    return fail(MPKX_ENOMEM, "fill_buffers: pass 2 allocation");
  }
  if (c.err) return fail(MPKX_ESTRUCT, "fill_buffers: a source value is never stored (schedule inconsistent)");
  for (int ph = 0; ph < nphase; ph++)
    if (comm[ph].cnt != cnt1[ph])
      return fail(MPKX_ESTRUCT, "fill_buffers: phase %d communication differs between passes", ph);

  mpkx_buffers **res = (mpkx_buffers **)calloc(npart, sizeof(mpkx_buffers *));
  if (!res) return fail(MPKX_ENOMEM, "fill_buffers: allocation");
  std::atomic<int> bad{0};
  int nthreads = (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
  std::atomic<long> next{0};
  auto worker = [&]() {
    std::vector<int> first;
    try {
      first.assign(spn, -1);
    } catch (const std::bad_alloc &) {
      bad = MPKX_ENOMEM;
      return;
    }
    for (;;) {
      long r = next.fetch_add(1);
      if (r >= npart || bad) break;
      mpkx_buffers *b = mpkx_new_bufs_internal(n, npart, nlevel, nphase, (int)r);
      if (!b) {
        bad = MPKX_ENOMEM;
        break;
      }
      res[r] = b;
      long cnt = 0, scnt = 0, mpc = 0;
      for (int ph = 0; ph <= nphase; ph++) {
        const PhaseComm &pc = comm[ph];
        int sc = 0, rc = 0;
        for (int p = 0; p < npart; p++) {
          int s_ = pc.cnt[(size_t)r * npart + p], r_ = pc.cnt[(size_t)p * npart + r];
          b->sendcounts[ph * npart + p] = s_;
          b->recvcounts[ph * npart + p] = r_;
          b->sdispls[ph * npart + p] = p ? b->sdispls[ph * npart + p - 1] + b->sendcounts[ph * npart + p - 1] : 0;
          b->rdispls[ph * npart + p] = p ? b->rdispls[ph * npart + p - 1] + b->recvcounts[ph * npart + p - 1] : 0;
          sc += s_;
          rc += r_;
        }
        b->scount[ph] = sc;
        b->rcount[ph] = rc;
        b->mcount[ph] = (int)mb[ph][r].size();
        b->rbuf_offsets[ph] = (int)cnt;
        cnt += rc;
        b->mbuf_offsets[ph] = (int)cnt;
        cnt += b->mcount[ph];
        b->sbuf_offsets[ph] = (int)scnt;
        scnt += sc;
        b->mptr_offsets[ph] = (int)mpc;
        mpc += b->mcount[ph] + 1;
      }
      if (cnt > INT_MAX || scnt > INT_MAX || mpc > INT_MAX) {
        bad = MPKX_EINVAL;
        break;
      }
      b->buf_count = (int)cnt;
      b->buf_scount = (int)scnt;
      b->mptr_count = (int)mpc;
      b->idx_buf = (long *)malloc(8 * (cnt + 1));
      b->idx_sbuf = (long *)malloc(8 * (scnt + 1));
      b->mptr_buf = (long *)malloc(8 * (mpc + 1));
      if (!b->idx_buf || !b->idx_sbuf || !b->mptr_buf) {
        bad = MPKX_ENOMEM;
        break;
      }
      long mcc = 0;
      for (int ph = 0; ph <= nphase; ph++) {
        long *ib = b->idx_buf + b->rbuf_offsets[ph];
        const auto &rv = comm[ph].recv[r];
        for (size_t t = 0; t < rv.size(); t++) ib[t] = rv[t];
        long *im = b->idx_buf + b->mbuf_offsets[ph];
        const auto &mv = mb[ph][r];
        long *mp = b->mptr_buf + b->mptr_offsets[ph];
        mp[0] = 0;
        for (size_t t = 0; t < mv.size(); t++) {
          im[t] = mv[t];
          int i = (int)(mv[t] % n);
          mp[t + 1] = mp[t] + ptr[i + 1] - ptr[i];
        }
        long *is = b->idx_sbuf + b->sbuf_offsets[ph];
        const auto &sv = comm[ph].send[r];
        for (size_t t = 0; t < sv.size(); t++) is[t] = sv[t];
        b->mcol_offsets[ph] = (int)mcc;
        mcc += mp[mv.size()];
      }
      if (mcc > INT_MAX) {
        bad = MPKX_EINVAL;
        break;
      }
      b->mcol_count = (int)mcc;
      b->mcol_buf = (long *)malloc(8 * (mcc + 1));
      b->mval_buf = val ? (double *)malloc(8 * (mcc + 1)) : nullptr;
      if (!b->mcol_buf || (val && !b->mval_buf)) {
        bad = MPKX_ENOMEM;
        break;
      }
      // first occurrence of every vv index in idx_buf (buffers.c:456-463)
      for (long i = 0; i < cnt; i++)
        if (first[b->idx_buf[i]] == -1) first[b->idx_buf[i]] = (int)i;
      for (int ph = 0; ph <= nphase && !bad; ph++) {
        const long *mp = b->mptr_buf + b->mptr_offsets[ph];
        const long *im = b->idx_buf + b->mbuf_offsets[ph];
        long *mc = b->mcol_buf + b->mcol_offsets[ph];
        double *mvl = b->mval_buf ? b->mval_buf + b->mcol_offsets[ph] : nullptr;
        for (int m = 0; m < b->mcount[ph]; m++) {
          long i = im[m] % n, level = im[m] / n;
          for (int j = ptr[i]; j < ptr[i + 1]; j++) {
            int f = first[col[j] + (long)n * (level - 1)];
            if (f < 0) bad = MPKX_ESTRUCT;
            mc[j - ptr[i] + mp[m]] = f;
            if (mvl) mvl[j - ptr[i] + mp[m]] = val[j];
          }
        }
      }
      for (long i = 0; i < scnt; i++) {
        int f = first[b->idx_sbuf[i]];
        if (f < 0) bad = MPKX_ESTRUCT;
        b->idx_sbuf[i] = f;
      }
      for (long i = 0; i < cnt; i++) first[b->idx_buf[i]] = -1;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; t++) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
  if (bad) {
    mpkx_del_bufs_array(res, npart);
    if (bad == MPKX_ESTRUCT) return fail(MPKX_ESTRUCT, "fill_buffers: a needed value is held nowhere in its rank");
    if (bad == MPKX_EINVAL) return fail(MPKX_EINVAL, "fill_buffers: a rank's buffer exceeds int range");
    return fail(MPKX_ENOMEM, "fill_buffers: allocation");
  }
  *out = res;
  return MPKX_OK;
}

// prep_mpk (mpk2/mpkprep.c:7-252).  The value array doubles as the
// bookkeeping of which (level, row) is computed by whom, as there.
int mpkx_prep_mpk(int n, const int *ptr, const int *col, int nlevel, int nphase,
                  const int *const *parts, const int *const *levels, const long *sptr,
                  const int *srow, const int *sdist, int npart, int *tcount, long **tidx_out,
                  long *ntask) {
  clear_error();
  if (n <= 0 || nlevel <= 0 || nphase < 0 || npart <= 0) return fail(MPKX_EINVAL, "prep_mpk: bad arguments");
  // level / skirt sanity checks, mpkprep.c:21-84
  for (int ph = 0; ph < nphase; ph++)
    for (int i = 0; i < n; i++) {
      if (ph == 0 && levels[0][i] < 0) return fail(MPKX_ESTRUCT, "level error at i=%d phase=%d", i, ph);
      if (ph > 0 && levels[ph - 1][i] > levels[ph][i])
        return fail(MPKX_ESTRUCT, "level error at i=%d phase=%d", i, ph);
    }
  for (int p = 0; p < npart; p++)
    for (long t = sptr[p]; t < sptr[p + 1]; t++) {
      int i = srow[t], d = sdist[t];
      if (nphase == 0 ? (nlevel - d < 0) : (levels[nphase - 1][i] > nlevel - d))
        return fail(MPKX_ESTRUCT, "skirt level error at i=%d p=%d", i, p);
    }
  long total_cap = (long)n * nlevel + (sptr[npart] + 1) * (long)(nlevel + 1);
  std::vector<double> mark;
  std::vector<long> tidx;
  MPKX_TRY_ALLOC(mark.assign((size_t)n * (nlevel + 1), -1.0); tidx.reserve(total_cap));
  for (int i = 0; i < n; i++) mark[i] = parts[0][i] * 100.0;
  std::vector<int> zero(n, 0);
  const int *prevl = zero.data();
  int prevlmin = 0;
  long err = 0;
  for (int ph = 0; ph < nphase; ph++) {
    const int *pl = parts[ph], *ll = levels[ph];
    int lmin = ll[0], lmax = ll[0];
    for (int i = 1; i < n; i++) lmin = std::min(lmin, ll[i]), lmax = std::max(lmax, ll[i]);
    lmax = std::min(lmax, nlevel);
    std::vector<long> size(npart, 0);
    for (int l = prevlmin + 1; l <= lmax; l++)
      for (int i = 0; i < n; i++)
        if (prevl[i] < l && l <= ll[i]) {
          if (mark[(size_t)l * n + i] >= 0) err++;
          for (int j = ptr[i]; j < ptr[i + 1]; j++) {
            int k = col[j];
            double m = mark[(size_t)(l - 1) * n + k];
            if (m < 0.0 || (pl[i] != pl[k] && ((int)m) % 100 == ph)) err++;
          }
          mark[(size_t)l * n + i] = pl[i] * 100.0 + ph;
          size[pl[i]]++;
        }
    std::vector<long> start(npart);
    long base = (long)tidx.size();
    for (int p = 0; p < npart; p++) {
      start[p] = base;
      base += size[p];
      tcount[ph * npart + p] = (int)size[p];
    }
    tidx.resize(base);
    for (int l = prevlmin + 1; l <= lmax; l++)
      for (int i = 0; i < n; i++)
        if (prevl[i] < l && l <= ll[i]) tidx[start[pl[i]]++] = (long)l * n + i;
    prevl = ll;
    prevlmin = lmin;
  }
  if (err) return fail(MPKX_ESTRUCT, "prep_mpk: structure error (%ld)", err);
  const int *pl0 = parts[0];
  for (int p = 0; p < npart; p++) {
    long cnt = 0;
    for (int l = prevlmin + 1; l <= nlevel; l++)
      for (long t = sptr[p]; t < sptr[p + 1]; t++) {
        int i = srow[t], d = sdist[t];
        if (prevl[i] < l && d >= 0 && l <= nlevel - d) {
          if (mark[(size_t)l * n + i] > 0.0) err++;
          for (int j = ptr[i]; j < ptr[i + 1]; j++)
            if (mark[(size_t)(l - 1) * n + col[j]] < 0.0) err++;
          mark[(size_t)l * n + i] = pl0[i] * 100.0 + nphase;
          tidx.push_back((long)l * n + i);
          cnt++;
        }
      }
    for (int l = prevlmin + 1; l <= nlevel; l++)
      for (long t = sptr[p]; t < sptr[p + 1]; t++) {
        int i = srow[t], d = sdist[t];
        if (prevl[i] < l && d >= 0 && l <= nlevel - d) mark[(size_t)l * n + i] = -0.1;
      }
    tcount[nphase * npart + p] = (int)cnt;
  }
  for (int i = 0; i < n; i++)
    if (mark[(size_t)nlevel * n + i] < -0.5) err++;
  if (err) return fail(MPKX_ESTRUCT, "prep_mpk: structure error (%ld)", err);
  long *t = (long *)malloc(sizeof(long) * (tidx.size() + 1));
  if (!t) return fail(MPKX_ENOMEM, "prep_mpk: allocation");
  memcpy(t, tidx.data(), sizeof(long) * tidx.size());
  *tidx_out = t;
  *ntask = (long)tidx.size();
  return MPKX_OK;
}

}  // extern "C"
