// spmv::L2GMap -- local-to-global index map with ghost entries (the
// interface of spmv/L2GMap.h:14-97), backed by the spmvx C ABI.
//
// One process drives one rank.  update / reverse_update work on DEVICE
// vectors (local_size + num_ghosts entries); what the other ranks would send
// comes from a host halo buffer (set_halo; zeros by default).  A rank of a
// larger decomposition is described with Decomposition: the exchange plan is
// then the one the reference builds with MPI.
#pragma once

#include <mpi.h>

#include <cstdint>
#include <vector>

#include "../spmv_amd.h"

namespace spmv
{

/// The other ranks of a decomposition driven by one process.
struct Decomposition
{
  std::vector<std::int64_t> local_sizes; ///< owned size of every rank
  int rank = 0;                          ///< this process's rank
  /// remote_ghosts[q]: sorted global indices of rank q's ghosts owned here
  std::vector<std::vector<std::int64_t>> remote_ghosts;
};

class L2GMap
{
public:
  /// L2GMap(comm, local_size, ghosts) (L2GMap.h:29-30).  comm must have one
  /// rank; ghosts must then be empty (a one-rank map has no owner for them).
  L2GMap(MPI_Comm comm, std::int64_t local_size, const std::vector<std::int64_t>& ghosts);
  /// One rank of `decomp` (comm has one rank; the others are virtual).
  L2GMap(MPI_Comm comm, std::int64_t local_size, const std::vector<std::int64_t>& ghosts,
         const Decomposition& decomp);
  ~L2GMap();
  L2GMap(const L2GMap&) = delete;
  L2GMap& operator=(const L2GMap&) = delete;

  std::int32_t local_size() const;
  std::int32_t num_ghosts() const;
  std::int64_t global_size() const;
  std::int64_t global_offset() const;
  std::int32_t global_to_local(std::int64_t i) const;
  const std::vector<std::int64_t>& ghosts() const { return _ghosts; }

  /// Ghost update on a device vector (L2GMap.cpp:126-150).
  template <typename T>
  void update(T* vec_data) const;
  /// Reverse update on a device vector (L2GMap.cpp:152-170); `incoming`
  /// (host, |indexbuf| values) is what the neighbours send back.
  template <typename T>
  void reverse_update(T* vec_data, const std::vector<T>& incoming = {}) const;

  /// Values the neighbours deliver into the ghost region (host).
  template <typename T>
  void set_halo(const std::vector<T>& halo);
  void clear_halo();

  spmvx_l2gmap handle() const { return _h; }

private:
  void build(MPI_Comm comm, std::int64_t local_size, const Decomposition& d);
  spmvx_l2gmap _h = nullptr;
  std::vector<std::int64_t> _ghosts;
};

} // namespace spmv
