// pdmpk.h -- host representation of pdmpk per-partition buffers (the fields
// pdmpk's Buffers::Dump writes, pdmpk/buffers.cc:194-206), shared by the
// prep (host_pdmpk.cpp) and the GPU exec (dev_pdmpk.hip).
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

#include "internal.h"

namespace mpkx {

// pdmpk/phase_descriptor.h:11-38: levels [bottom, top) of a phase; levels
// below `mid` wait for their incoming data.
struct PdPhaseDescriptor {
  short bottom = -1, mid = 0, top = -1;
};

struct PdBuffers {
  size_t max_sbuf_size = 0;
  int mbuf_idx = 0;                          // mbuf size
  std::vector<size_t> mbuf_begin;            // mbuf.phase_begin (per batch)
  std::vector<int> home_idcs, original_idcs;
  std::vector<PdPhaseDescriptor> pds;
  std::vector<int> res_mbuf_idcs, res_original_idcs;  // Results
  std::vector<int> sendcounts, sdispls, recvcounts, rdispls;  // npart per batch
  std::vector<int> sbuf_idcs;
  std::vector<size_t> sbuf_begin;
  std::vector<std::pair<int, int>> init_idcs;  // (source mbuf idx, target mbuf idx)
  std::vector<size_t> init_begin;
  std::vector<int> mptr;
  std::vector<size_t> mptr_begin;
  std::vector<int> mcol;
  std::vector<double> mval;
};

}  // namespace mpkx

struct mpkx_pd;
namespace mpkx {
const std::vector<PdBuffers> &pd_buffers(const mpkx_pd *pd);
int pd_n(const mpkx_pd *pd);
}  // namespace mpkx
