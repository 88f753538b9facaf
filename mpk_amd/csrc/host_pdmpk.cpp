// host_pdmpk.cpp -- the pdmpk "partial diamond" schedule and its buffers.
//
// Restates pdmpk's prep (pdmpk/comm_comp_patterns.cc, pdmpk_buffers.cc,
// buffers.cc, mpi_buffers.cc, mcsr.cc, results.cc, phase_descriptor.cc):
//   - vertices advance one level per batch; a row may be completed over
//     several batches from partial sums ("partials": the edges already
//     added), a partial sum being carried to the next batch through
//     `init_idcs` (or communicated when it lives in another partition);
//   - after each phase the partitions are redrawn and relabelled by LAPJV
//     so the new labels keep most values where they are;
//   - final values are sent home at the end (SendHome / Epilogue).
// The reference redraws partitions with METIS on level-based edge weights
// (pdmpk_buffers.cc:183-201).  METIS is not available, so here the caller
// supplies the partitions: `part0` for the start and a list of candidate
// partitions used, in turn, wherever the reference calls METIS; the rest of
// NewPartitionLabels (MetisFixVertex, OptimizeLabels with LAPJV) is kept.
//
// Output per partition: the Buffers fields written by Buffers::Dump
// (pdmpk/buffers.cc:194-206) and the Results index (results.cc:31-40).
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "internal.h"
#include "pdmpk.h"

using namespace mpkx;

namespace {

enum CommType { kMcol = 0, kInitIdcs = 1, kFinished = 2 };

// PdmpkBuffers (pdmpk/pdmpk_buffers.cc): partial-edge bits, partitions and
// levels of every vertex.
struct PdState {
  const int *ptr, *col;
  int n;
  std::vector<char> partials;
  std::vector<int> partitions;
  std::vector<short> levels;

  bool CanAdd(int idx, int lbelow, int t) const {  // pdmpk_buffers.cc:57-65
    int j = col[t];
    return !partials[t] && partitions[idx] == partitions[j] && levels[j] >= lbelow;
  }
  bool PartialIsFull(int idx) const {
    for (int t = ptr[idx]; t < ptr[idx + 1]; t++)
      if (!partials[t]) return false;
    return true;
  }
  bool PartialIsEmpty(int idx) const {
    for (int t = ptr[idx]; t < ptr[idx + 1]; t++)
      if (partials[t]) return false;
    return true;
  }
  void IncLevel(int idx) {  // pdmpk_buffers.cc:67-72
    if (PartialIsFull(idx)) {
      levels[idx]++;
      for (int t = ptr[idx]; t < ptr[idx + 1]; t++) partials[t] = 0;
    }
  }
  short MinLevel() const { return *std::min_element(levels.begin(), levels.end()); }
  size_t ExactLevelSum() const {  // pdmpk_buffers.cc:30-46
    size_t sum = 0;
    for (int i = 0; i < n; i++) {
      sum += (size_t)levels[i] * (ptr[i + 1] - ptr[i]);
      for (int t = ptr[i]; t < ptr[i + 1]; t++) sum += partials[t] ? 1 : 0;
    }
    return sum;
  }
  bool IsFinished(int nlevel) const {
    for (short l : levels)
      if (l != nlevel) return false;
    return true;
  }
};

class Patterns {
 public:
  Patterns(int n, const int *ptr, const int *col, const double *val, int npart, int nlevel,
           const int *part0, int nrepart, const int *const *reparts, int mirror)
      : n_(n), ptr_(ptr), col_(col), val_(val), np_(npart), nl_(nlevel), nrep_(nrepart),
        reparts_(reparts), mirror_(mirror) {
    bufs_.resize(npart);
    st_.ptr = ptr;
    st_.col = col;
    st_.n = n;
    st_.partials.assign(ptr[n], 0);
    st_.partitions.assign(part0, part0 + n);
    st_.levels.assign(n, 0);
    sp_part_.assign((size_t)n * (nlevel + 1), -1);
    sp_idx_.assign((size_t)n * (nlevel + 1), -1);
  }

  int run() {
    // constructor of pdmpk/comm_comp_patterns.cc:26-64
    history_.push_back(st_.partitions);
    home_ = st_.partitions;
    for (int idx = 0; idx < n_; idx++) {
      int part = st_.partitions[idx];
      bufs_[part].home_idcs.push_back(0);
      bufs_[part].original_idcs.push_back(idx);
      finalize(idx, 0, part);
    }
    batch_ = -1;
    proc_phase(0);
    for (auto &b : bufs_) update_mid(b.pds.back(), -1);
    int rc = all_phases();
    if (rc) return rc;
    epilogue();
    return check();
  }

  std::vector<PdBuffers> bufs_;

 private:
  int n_;
  const int *ptr_, *col_;
  const double *val_;
  int np_, nl_, nrep_;
  const int *const *reparts_;
  int mirror_;
  int batch_ = -1;
  int nrepart_used_ = 0;
  PdState st_;
  std::vector<int> sp_part_, sp_idx_;  // store_part[(idx, lvl)] = (part, mbuf idx)
  std::vector<int> home_;
  std::vector<std::vector<int>> history_;
  // The batch's communication (the reference's comm_dict: a std::map of
  // (src, tgt) to a std::map of (src_mbuf_idx, type) to a std::set of ints) as flat
  // records; sorted by (src, tgt, src_mbuf_idx, type, value) with exact
  // duplicates removed they group and iterate in the nested containers'
  // order, without a tree node per insert.
  struct CommRec {
    int src, tgt, m, type, v;
    bool operator<(const CommRec &o) const {
      return std::tie(src, tgt, m, type, v) < std::tie(o.src, o.tgt, o.m, o.type, o.v);
    }
    bool operator==(const CommRec &o) const {
      return src == o.src && tgt == o.tgt && m == o.m && type == o.type && v == o.v;
    }
  };
  std::vector<CommRec> comm_;
  void comm_add(int src, int tgt, int m, int type, int v) { comm_.push_back({src, tgt, m, type, v}); }

  // level-major: the neighbours of a vertex at one level are close in memory
  size_t key(int idx, int lvl) const { return (size_t)lvl * n_ + idx; }
  bool sp_has(int idx, int lvl) const { return sp_part_[key(idx, lvl)] >= 0; }

  static void update_bottom(PdPhaseDescriptor &pd, short l) {  // phase_descriptor.cc:10-15
    if (pd.bottom == -1) {
      pd.bottom = l;
      pd.mid = l;
    }
  }
  static void update_mid(PdPhaseDescriptor &pd, short l) { pd.mid = pd.mid <= l ? l + 1 : pd.mid; }
  static void update_top(PdPhaseDescriptor &pd, short l) { pd.top = pd.top <= l ? l + 1 : pd.top; }

  void finalize(int idx, int lvl, int part) {  // FinalizeVertex, :666-671
    int &m = bufs_[part].mbuf_idx;
    sp_part_[key(idx, lvl)] = part;
    sp_idx_[key(idx, lvl)] = m;
    m++;
  }

  void send_home(int idx) {  // :666-675 (SendHome)
    if (!sp_has(idx, nl_)) return;
    int tgt = home_[idx];
    int src = sp_part_[key(idx, nl_)], m = sp_idx_[key(idx, nl_)];
    if (src != tgt) comm_add(src, tgt, m, kFinished, idx);
  }

  void pre_batch() {  // :481-495 PreBatch + Buffers::PreBatch (buffers.cc:30-34)
    batch_++;
    for (auto &b : bufs_) {
      b.sendcounts.resize(b.sendcounts.size() + np_, 0);
      b.recvcounts.resize(b.recvcounts.size() + np_, 0);
      b.sdispls.resize(b.sdispls.size() + np_, 0);
      b.rdispls.resize(b.rdispls.size() + np_, 0);
      b.sbuf_begin.push_back(b.sbuf_idcs.size());
      b.init_begin.push_back(b.init_idcs.size());
      b.mptr_begin.push_back(b.mptr.size());
      b.mbuf_begin.push_back((size_t)b.mbuf_idx);
    }
    for (int idx = 0; idx < n_; idx++)
      if (st_.levels[idx] == nl_) send_home(idx);
  }

  size_t sbuf_size(const PdBuffers &b, int batch) const {
    size_t i = (size_t)np_ * batch + np_ - 1;
    return (size_t)b.sendcounts[i] + b.sdispls[i];
  }
  size_t rbuf_size(const PdBuffers &b, int batch) const {
    size_t i = (size_t)np_ * batch + np_ - 1;
    return (size_t)b.recvcounts[i] + b.rdispls[i];
  }

  void post_batch(short lbelow) {  // :520-545 PostBatch
    std::sort(comm_.begin(), comm_.end());
    comm_.erase(std::unique(comm_.begin(), comm_.end()), comm_.end());
    // (src, tgt) groups [g0, g1): each distinct (m, type) is one value sent
    std::vector<std::pair<size_t, size_t>> groups;
    for (size_t a = 0; a < comm_.size();) {
      size_t b = a;
      while (b < comm_.size() && comm_[b].src == comm_[a].src && comm_[b].tgt == comm_[a].tgt) b++;
      groups.push_back({a, b});
      a = b;
    }
    for (auto &g : groups) {  // UpdateMpiCountBuffers, :614-625
      int src = comm_[g.first].src, tgt = comm_[g.first].tgt;
      int size = 0;
      for (size_t r = g.first; r < g.second; r++)
        if (r == g.first || comm_[r].m != comm_[r - 1].m || comm_[r].type != comm_[r - 1].type) size++;
      bufs_[tgt].recvcounts[(size_t)batch_ * np_ + src] += size;
      bufs_[src].sendcounts[(size_t)batch_ * np_ + tgt] += size;
    }
    for (auto &b : bufs_) {  // Buffers::PostBatch, buffers.cc:36-50
      size_t off = b.sendcounts.size() - np_;
      b.sdispls[off] = 0;
      b.rdispls[off] = 0;
      for (int i = 1; i < np_; i++) {
        b.sdispls[off + i] = b.sdispls[off + i - 1] + b.sendcounts[off + i - 1];
        b.rdispls[off + i] = b.rdispls[off + i - 1] + b.recvcounts[off + i - 1];
      }
      size_t ss = sbuf_size(b, batch_);
      if (b.max_sbuf_size < ss) b.max_sbuf_size = ss;
      b.sbuf_idcs.resize(b.sbuf_idcs.size() + ss, 0);
      b.mbuf_idx += (int)rbuf_size(b, batch_);
    }
    std::vector<char> needs_wait(np_, 0);
    for (auto &g : groups) {
      needs_wait[comm_[g.first].tgt] = 1;
      proc_comm_dict(g.first, g.second);
    }
    comm_.clear();
    for (int i = 0; i < np_; i++) {
      auto &pd = bufs_[i].pds.back();
      update_bottom(pd, lbelow);
      update_top(pd, lbelow);
      if (needs_wait[i]) update_mid(pd, lbelow);
      auto &ii = bufs_[i].init_idcs;  // MpiBuffers::SortInitIdcs
      std::sort(ii.begin() + bufs_[i].init_begin.back(), ii.end(),
                [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.second < b.second; });
    }
  }

  // One (src, tgt) group of records [g0, g1), sorted: every distinct
  // (src_mbuf_idx, type) is one sent value; its records' v ascending.
  void proc_comm_dict(size_t g0, size_t g1) {  // :627-664
    const int src = comm_[g0].src, tgt = comm_[g0].tgt;
    PdBuffers &sb = bufs_[src];
    PdBuffers &tb = bufs_[tgt];
    size_t send_base = sb.sbuf_begin[batch_] + sb.sdispls[(size_t)batch_ * np_ + tgt];
    size_t recv_base = tb.mbuf_begin[batch_] + tb.mptr.size() - tb.mptr_begin[batch_] +
                       tb.rdispls[(size_t)batch_ * np_ + src];
    size_t idx = 0;
    for (size_t e0 = g0; e0 < g1;) {
      size_t e1 = e0;
      while (e1 < g1 && comm_[e1].m == comm_[e0].m && comm_[e1].type == comm_[e0].type) e1++;
      int src_idx = (int)(recv_base + idx);
      if (comm_[e0].type == kMcol) {
        for (size_t r = e0; r < e1; r++) tb.mcol[comm_[r].v] = src_idx;
      } else if (comm_[e0].type == kInitIdcs) {
        tb.init_idcs.push_back({src_idx, comm_[e0].v});
      } else {
        int orig = comm_[e0].v;
        sp_part_[key(orig, nl_)] = tgt;
        sp_idx_[key(orig, nl_)] = src_idx;
      }
      sb.sbuf_idcs[send_base + idx] = comm_[e0].m;
      idx++;
      e0 = e1;
    }
  }

  void add_to_init(int idx, int level) {  // :571-584
    int sp = sp_part_[key(idx, level)], si = sp_idx_[key(idx, level)];
    int tp = st_.partitions[idx];
    int ti = bufs_[tp].mbuf_idx;
    if (sp != tp)
      comm_add(sp, tp, si, kInitIdcs, ti);
    else
      bufs_[tp].init_idcs.push_back({si, ti});
  }

  void proc_adjacent(int idx, int lbelow, int t) {  // :586-607
    st_.partials[t] = 1;
    int j = col_[t];
    int cp = st_.partitions[idx];
    int sp = sp_part_[key(j, lbelow)], si = sp_idx_[key(j, lbelow)];
    PdBuffers &b = bufs_[cp];
    b.mval.push_back(val_[t]);
    if (cp == sp) {
      b.mcol.push_back(si);
    } else {
      int ti = (int)b.mcol.size();
      comm_add(sp, cp, si, kMcol, ti);
      b.mcol.push_back(-1);
    }
  }

  bool proc_vertex(int idx, int lbelow) {  // :547-569
    bool ret = false;
    int cp = st_.partitions[idx];
    bool send_partial = !st_.PartialIsEmpty(idx);
    for (int t = ptr_[idx]; t < ptr_[idx + 1]; t++)
      if (st_.CanAdd(idx, lbelow, t)) {
        if (!ret) bufs_[cp].mptr.push_back((int)bufs_[cp].mcol.size());
        proc_adjacent(idx, lbelow, t);
        ret = true;
      }
    if (ret) {
      if (send_partial) add_to_init(idx, lbelow + 1);
      st_.IncLevel(idx);
      finalize(idx, lbelow + 1, cp);
    }
    return ret;
  }

  void proc_phase(short min_level) {  // :481-518 ProcPhase
    for (auto &b : bufs_) b.pds.push_back(PdPhaseDescriptor());
    bool active = true;
    for (int lbelow = min_level; active && lbelow < nl_; lbelow++) {
      pre_batch();
      active = false;
      for (int idx = 0; idx < n_; idx++)
        if (st_.levels[idx] == lbelow && proc_vertex(idx, lbelow)) active = true;
      post_batch((short)lbelow);
    }
  }

  // OptimizeLabels (:415-431) + OptimizeVertex (:433-461) +
  // FindLabelPermutation (:463-479) on a copy of the state.
  void optimize_labels(short min_level) {
    PdState cnt = st_;
    // (s, t) -> the distinct source slots: only the counts are used, so
    // keys (s*np + t, slot) sorted and made unique
    std::vector<uint64_t> table;
    auto tab = [&](int s, int t, int slot) {
      table.push_back(((uint64_t)((long)s * np_ + t) << 32) | (uint32_t)slot);
    };
    bool active = true;
    for (int lbelow = min_level; active && lbelow < nl_; lbelow++) {
      active = false;
      for (int idx = 0; idx < n_; idx++) {
        if (cnt.levels[idx] != lbelow) continue;
        bool ret = false;
        int tp = cnt.partitions[idx];
        bool send_partial = !cnt.PartialIsEmpty(idx);
        for (int t = ptr_[idx]; t < ptr_[idx + 1]; t++)
          if (cnt.CanAdd(idx, lbelow, t)) {
            cnt.partials[t] = 1;
            int j = col_[t];
            if (sp_has(j, lbelow)) {
              tab(sp_part_[key(j, lbelow)], tp, sp_idx_[key(j, lbelow)]);
              ret = true;
            }
          }
        if (ret) {
          if (send_partial) tab(sp_part_[key(idx, lbelow + 1)], tp, sp_idx_[key(idx, lbelow + 1)]);
          cnt.IncLevel(idx);
          active = true;
        }
      }
    }
    std::sort(table.begin(), table.end());
    table.erase(std::unique(table.begin(), table.end()), table.end());
    std::vector<int> sums((size_t)np_ * np_, 0);
    for (uint64_t k : table) sums[k >> 32]++;
    int mn = *std::min_element(sums.begin(), sums.end());
    for (auto &v : sums) v -= mn;
    std::vector<int> perm(np_);
    mpkx_lapjv(sums.data(), np_, perm.data());
    for (auto &p : st_.partitions) p = perm[p];
  }

  // NewPartitionLabels (:396-413): the partition source stands in for
  // METIS_PartGraphKway on the weighted graph.
  void new_partition_labels(short min_level) {
    const int *src = nrep_ > 0 ? reparts_[nrepart_used_ % nrep_] : history_[0].data();
    nrepart_used_++;
    st_.partitions.assign(src, src + n_);
    for (int idx = 0; idx < n_; idx++)
      if (st_.levels[idx] == min_level)  // MetisFixVertex, pdmpk_buffers.cc:204-211
        for (int t = ptr_[idx]; t < ptr_[idx + 1]; t++)
          if (!st_.partials[t]) st_.partitions[col_[t]] = st_.partitions[idx];
    optimize_labels(min_level);
    history_.push_back(st_.partitions);
  }

  // ProcAllPhases0..5 (:245-394): how the next phase's partition is chosen.
  int all_phases() {
    size_t old_sum = 0;
    bool finished = st_.IsFinished(nl_);
    auto stuck = [&]() { return fail(MPKX_ESTRUCT, "pdmpk prep: could not finish (levels stopped growing)"); };
    long guard = 0;
    while (!finished) {
      if (++guard > 100000) return stuck();
      short mn = st_.MinLevel();
      size_t sum = st_.ExactLevelSum();
      switch (mirror_) {
        case 0:
          if (old_sum == sum) return stuck();
          old_sum = sum;
          new_partition_labels(mn);
          break;
        case 1:
        case 5: {
          bool redraw = mirror_ == 1 ? mn < nl_ / 2 : 2 * mn < 3 * nl_;
          if (history_.empty()) return stuck();
          if (redraw) {
            if (old_sum == sum) return stuck();
            old_sum = sum;
            new_partition_labels(mn);
          } else {
            st_.partitions = history_.back();
            history_.pop_back();
          }
        } break;
        case 2:
          if (history_.empty()) return stuck();
          if (mn == 0) {
            if (old_sum == sum) return stuck();
            old_sum = sum;
            new_partition_labels(mn);
          } else {
            st_.partitions = history_[(size_t)batch_ % history_.size()];
          }
          break;
        case 3:
          if (history_.empty()) return stuck();
          if (mn < nl_ / 2) {
            new_partition_labels(mn);
          } else {
            st_.partitions = history_.back();
            history_.pop_back();
          }
          break;
        case 4:
          if (history_.empty()) return stuck();
          if (mn == 0)
            new_partition_labels(mn);
          else
            st_.partitions = history_[(size_t)batch_ % history_.size()];
          break;
        default:
          return fail(MPKX_EINVAL, "pdmpk prep: mirror method %d outside 0..5", mirror_);
      }
      proc_phase(mn);
      finished = st_.IsFinished(nl_);
    }
    return MPKX_OK;
  }

  void epilogue() {  // :66-96
    for (int idx = 0; idx < n_; idx++) {
      if (home_[idx] != sp_part_[key(idx, nl_)]) {
        for (auto &b : bufs_) b.pds.push_back(PdPhaseDescriptor());
        pre_batch();
        post_batch((short)(nl_ - 1));
        break;
      }
    }
    for (auto &b : bufs_) {
      b.mptr_begin.push_back(b.mptr.size());
      b.mptr.push_back((int)b.mcol.size());
      b.init_begin.push_back(b.init_idcs.size());
    }
    for (int i = 0; i < n_; i++) {
      int part = sp_part_[key(i, nl_)], m = sp_idx_[key(i, nl_)];
      int m0 = sp_idx_[key(i, 0)];
      bufs_[part].home_idcs[m0] = m;
      bufs_[part].res_original_idcs.push_back(i);
      bufs_[part].res_mbuf_idcs.push_back(m);
    }
  }

  // DbgAsserts (:700-745): every vertex reached nlevel, mbuf = mptr + rbuf
  // per batch, the phase descriptors cover every batch, final values home.
  int check() const {
    for (short l : st_.levels)
      if (l != nl_) return fail(MPKX_ESTRUCT, "pdmpk prep: a vertex did not reach nlevel");
    for (int i = 0; i < n_; i++)
      if (sp_part_[key(i, nl_)] != sp_part_[key(i, 0)])
        return fail(MPKX_ESTRUCT, "pdmpk prep: a final value is not home");
    for (auto &b : bufs_) {
      size_t nb = b.mbuf_begin.size();
      for (size_t bt = 1; bt < nb; bt++) {
        size_t mbd = b.mbuf_begin[bt] - b.mbuf_begin[bt - 1];
        size_t mpd = b.mptr_begin[bt] - b.mptr_begin[bt - 1];
        if (mbd != mpd + rbuf_size(b, (int)bt - 1))
          return fail(MPKX_ESTRUCT, "pdmpk prep: mbuf != mptr + rbuf at batch %zu", bt);
      }
      size_t sum = 0;
      for (auto &pd : b.pds) {
        int d = pd.top - pd.bottom;
        sum += d ? d : 0;
      }
      if (sum != nb) return fail(MPKX_ESTRUCT, "pdmpk prep: phase descriptors cover %zu of %zu batches", sum, nb);
      for (int c : b.mcol)
        if (c < 0 || c >= b.mbuf_idx) return fail(MPKX_ESTRUCT, "pdmpk prep: unresolved matrix column");
    }
    return MPKX_OK;
  }
};

template <class T>
void dump_vec(FILE *f, const std::vector<T> &v) {  // Utils::DumpVec, pdmpk/utils.hpp:19-24
  size_t s = v.size();
  fwrite(&s, sizeof s, 1, f);
  if (s) fwrite(v.data(), sizeof(T), s, f);
}

template <class T>
bool load_vec(FILE *f, std::vector<T> &v) {
  size_t s = 0;
  if (fread(&s, sizeof s, 1, f) != 1 || s > ((size_t)1 << 40)) return false;
  v.resize(s);
  return s == 0 || fread(v.data(), sizeof(T), s, f) == s;
}

}  // namespace

struct mpkx_pd {
  int n = 0, npart = 0, nlevel = 0;
  std::vector<PdBuffers> bufs;
};

namespace mpkx {
const std::vector<PdBuffers> &pd_buffers(const mpkx_pd *pd) { return pd->bufs; }
int pd_n(const mpkx_pd *pd) { return pd->n; }
}  // namespace mpkx

extern "C" {

int mpkx_pd_prep(mpkx_pd **out, int n, const int *ptr, const int *col, const double *val, int npart,
                 int nlevel, const int *part0, int nrepart, const int *const *reparts, int mirror) {
  clear_error();
  if (!out || n <= 0 || !ptr || !col || !val || npart <= 0 || nlevel <= 0 || nlevel > SHRT_MAX - 1 || !part0 ||
      nrepart < 0 || (nrepart > 0 && !reparts))
    return fail(MPKX_EINVAL, "pd_prep: bad arguments");
  for (int i = 0; i < n; i++)
    if (part0[i] < 0 || part0[i] >= npart) return fail(MPKX_EINVAL, "pd_prep: label out of range");
  for (int r = 0; r < nrepart; r++)
    for (int i = 0; i < n; i++)
      if (reparts[r][i] < 0 || reparts[r][i] >= npart) return fail(MPKX_EINVAL, "pd_prep: label out of range");
  auto *pd = new (std::nothrow) mpkx_pd;
  if (!pd) return fail(MPKX_ENOMEM, "pd_prep");
  std::unique_ptr<mpkx_pd> guard(pd);
  pd->n = n;
  pd->npart = npart;
  pd->nlevel = nlevel;
  try {
    Patterns pat(n, ptr, col, val, npart, nlevel, part0, nrepart, reparts, mirror);
    int rc = pat.run();
    if (rc) return rc;
    pd->bufs = std::move(pat.bufs_);
  } catch (const std::bad_alloc &) {
    return fail(MPKX_ENOMEM, "pd_prep: host allocation");
  }
  *out = guard.release();
  return MPKX_OK;
}

void mpkx_pd_destroy(mpkx_pd *pd) { delete pd; }

int mpkx_pd_info(const mpkx_pd *pd, int *npart, int *nbatch, long *mbuf_total, long *mcol_total) {
  clear_error();
  if (!pd) return fail(MPKX_EINVAL, "pd_info: null");
  long mb = 0, mc = 0;
  for (auto &b : pd->bufs) {
    mb += b.mbuf_idx;
    mc += (long)b.mcol.size();
  }
  if (npart) *npart = pd->npart;
  if (nbatch) *nbatch = pd->bufs.empty() ? 0 : (int)pd->bufs[0].mbuf_begin.size();
  if (mbuf_total) *mbuf_total = mb;
  if (mcol_total) *mcol_total = mc;
  return MPKX_OK;
}

// Buffers::Dump (pdmpk/buffers.cc:194-206) -> <prefix>-r<rank>.buf.bin and
// Results::Dump (pdmpk/results.cc:36-40) -> <prefix>-r<rank>.results.bin.
int mpkx_pd_dump(const mpkx_pd *pd, const char *prefix) {
  clear_error();
  if (!pd || !prefix) return fail(MPKX_EINVAL, "pd_dump: bad arguments");
  for (int r = 0; r < pd->npart; r++) {
    const PdBuffers &b = pd->bufs[r];
    std::string base = std::string(prefix) + "-r" + std::to_string(r);
    FILE *f = fopen((base + ".buf.bin").c_str(), "wb");
    if (!f) return fail(MPKX_EIO, "cannot open %s.buf.bin", base.c_str());
    fwrite(&b.max_sbuf_size, sizeof b.max_sbuf_size, 1, f);
    fwrite(&b.mbuf_idx, sizeof b.mbuf_idx, 1, f);
    dump_vec(f, b.mbuf_begin);
    dump_vec(f, b.home_idcs);
    dump_vec(f, b.original_idcs);
    dump_vec(f, b.pds);
    dump_vec(f, b.res_mbuf_idcs);
    dump_vec(f, b.sendcounts);
    dump_vec(f, b.sdispls);
    dump_vec(f, b.recvcounts);
    dump_vec(f, b.rdispls);
    dump_vec(f, b.sbuf_idcs);
    dump_vec(f, b.sbuf_begin);
    dump_vec(f, b.init_idcs);
    dump_vec(f, b.init_begin);
    dump_vec(f, b.mptr);
    dump_vec(f, b.mptr_begin);
    dump_vec(f, b.mcol);
    dump_vec(f, b.mval);
    fclose(f);
    FILE *g = fopen((base + ".results.bin").c_str(), "wb");
    if (!g) return fail(MPKX_EIO, "cannot open %s.results.bin", base.c_str());
    dump_vec(g, b.res_original_idcs);
    dump_vec(g, std::vector<double>());
    fclose(g);
  }
  return MPKX_OK;
}

// Buffers::Load + Results::Load (pdmpk/buffers.cc:208-224, results.cc:42-46).
int mpkx_pd_load(mpkx_pd **out, const char *prefix, int npart, int n, int nlevel) {
  clear_error();
  if (!out || !prefix || npart <= 0 || n <= 0) return fail(MPKX_EINVAL, "pd_load: bad arguments");
  auto *pd = new (std::nothrow) mpkx_pd;
  if (!pd) return fail(MPKX_ENOMEM, "pd_load");
  std::unique_ptr<mpkx_pd> guard(pd);
  pd->n = n;
  pd->npart = npart;
  pd->nlevel = nlevel;
  pd->bufs.resize(npart);
  for (int r = 0; r < npart; r++) {
    PdBuffers &b = pd->bufs[r];
    std::string base = std::string(prefix) + "-r" + std::to_string(r);
    FILE *f = fopen((base + ".buf.bin").c_str(), "rb");
    if (!f) return fail(MPKX_EIO, "cannot open %s.buf.bin", base.c_str());
    bool ok = fread(&b.max_sbuf_size, sizeof b.max_sbuf_size, 1, f) == 1 &&
              fread(&b.mbuf_idx, sizeof b.mbuf_idx, 1, f) == 1 && load_vec(f, b.mbuf_begin) &&
              load_vec(f, b.home_idcs) && load_vec(f, b.original_idcs) && load_vec(f, b.pds) &&
              load_vec(f, b.res_mbuf_idcs) && load_vec(f, b.sendcounts) && load_vec(f, b.sdispls) &&
              load_vec(f, b.recvcounts) && load_vec(f, b.rdispls) && load_vec(f, b.sbuf_idcs) &&
              load_vec(f, b.sbuf_begin) && load_vec(f, b.init_idcs) && load_vec(f, b.init_begin) &&
              load_vec(f, b.mptr) && load_vec(f, b.mptr_begin) && load_vec(f, b.mcol) && load_vec(f, b.mval);
    fclose(f);
    if (!ok) return fail(MPKX_EIO, "%s.buf.bin: truncated", base.c_str());
    FILE *g = fopen((base + ".results.bin").c_str(), "rb");
    if (!g) return fail(MPKX_EIO, "cannot open %s.results.bin", base.c_str());
    std::vector<double> vals;
    ok = load_vec(g, b.res_original_idcs) && load_vec(g, vals);
    fclose(g);
    if (!ok || b.res_original_idcs.size() != b.res_mbuf_idcs.size())
      return fail(MPKX_EIO, "%s.results.bin: bad", base.c_str());
  }
  *out = guard.release();
  return MPKX_OK;
}

}  // extern "C"
