// L2GMap: exchange plan (spmv/L2GMap.cpp:15-122), accessors (:172-199) and
// the single-process ghost transport of update / reverse_update
// (:126-170): device pack of the owned entries the neighbours ghost, device
// unpack of the halo into the contiguous ghost region.
#include <algorithm>
#include <atomic>
#include <cstring>

#include <type_traits>

#include "handles.hpp"

namespace spmvx
{
namespace
{

// send[i] = v[indexbuf[i]] (L2GMap.cpp:137-139) and
// v[local_size + g] = inbox[g] (what Neighbor_alltoallv lands, :144-147).
template <typename T>
__global__ void halo_pack_unpack(const int32_t* __restrict__ indexbuf, int64_t nidx,
                                 T* __restrict__ v, T* __restrict__ send,
                                 const T* __restrict__ inbox, int64_t local_size,
                                 int64_t nghost)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t n = nidx > nghost ? nidx : nghost;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
  {
    if (i < nidx)
      send[i] = v[indexbuf[i]];
    if (i < nghost)
      v[local_size + i] = inbox ? inbox[i] : T(0);
  }
}

// v[target[t]] += in[pos[j]] for j in rev_ptr[t]..rev_ptr[t+1], ascending:
// the sequential loop of L2GMap.cpp:168-169 with repeated indices, without
// atomics and in the same order.
template <typename T>
__global__ void reverse_scatter_add(const int32_t* __restrict__ target,
                                    const int32_t* __restrict__ ptr,
                                    const int32_t* __restrict__ pos, int64_t ntarget,
                                    const T* __restrict__ in, T* __restrict__ v)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < ntarget; t += stride)
  {
    T acc = v[target[t]];
    for (int32_t j = ptr[t]; j < ptr[t + 1]; ++j)
      acc = acc + in[pos[j]];
    v[target[t]] = acc;
  }
}

int grid_for(int64_t n)
{
  int64_t g = (n + 255) / 256;
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

size_t vec_elem(int32_t dtype)
{
  if (dtype == SPMVX_F64) return 8;
  if (dtype == SPMVX_F32) return 4;
  if (dtype == SPMVX_C64) return 16;
  if (dtype == SPMVX_C32) return 8;
  throw Error(SPMVX_ERR_ARG, "vectors must be F64, F32, C64 or C32");
}

// Calls f with a null pointer of the element type of dtype.
template <typename F>
void with_vec_type(int32_t dtype, F&& f)
{
  switch (dtype)
  {
  case SPMVX_F64: f(static_cast<double*>(nullptr)); break;
  case SPMVX_F32: f(static_cast<float*>(nullptr)); break;
  case SPMVX_C64: f(static_cast<cplx<double>*>(nullptr)); break;
  case SPMVX_C32: f(static_cast<cplx<float>*>(nullptr)); break;
  default: throw Error(SPMVX_ERR_ARG, "vectors must be F64, F32, C64 or C32");
  }
}

} // namespace

void l2gmap_prepare_device(spmvx_l2gmap_s* m)
{
  if (m->device_ready)
    return;
  const size_t nidx = m->indexbuf.size();
  m->d_indexbuf.ensure(4 * std::max<size_t>(nidx, 1));
  if (nidx)
    SPMVX_HIP(hipMemcpy(m->d_indexbuf.ptr, m->indexbuf.data(), 4 * nidx, hipMemcpyHostToDevice));
  m->d_send.ensure(16 * std::max<size_t>(nidx, 1)); // up to complex double
  const size_t nt = m->rev_target.size();
  m->d_rev_target.ensure(4 * std::max<size_t>(nt, 1));
  m->d_rev_ptr.ensure(4 * (nt + 1));
  m->d_rev_pos.ensure(4 * std::max<size_t>(m->rev_pos.size(), 1));
  if (nt)
  {
    SPMVX_HIP(hipMemcpy(m->d_rev_target.ptr, m->rev_target.data(), 4 * nt, hipMemcpyHostToDevice));
    SPMVX_HIP(hipMemcpy(m->d_rev_pos.ptr, m->rev_pos.data(), 4 * m->rev_pos.size(), hipMemcpyHostToDevice));
  }
  SPMVX_HIP(hipMemcpy(m->d_rev_ptr.ptr, m->rev_ptr.data(), 4 * (nt + 1), hipMemcpyHostToDevice));
  m->device_ready = true;
}

void l2gmap_update_async(spmvx_l2gmap_s* m, int32_t dtype, void* vec, hipStream_t s)
{
  l2gmap_prepare_device(m);
  const int64_t nidx = static_cast<int64_t>(m->indexbuf.size());
  const int64_t ng = static_cast<int64_t>(m->ghosts.size());
  if (nidx == 0 && ng == 0)
    return;
  (void)vec_elem(dtype); // rejects an unknown dtype
  // receive: the unpack kernel reads the device copy of the halo (what the
  // neighbours send) made by set_halo
  const void* inbox = nullptr;
  if (ng > 0 && m->halo_dtype >= 0)
  {
    SPMVX_REQUIRE(m->halo_dtype == dtype, "halo dtype differs from the vector dtype");
    inbox = m->d_halo.ptr;
  }
  const int g = grid_for(std::max(nidx, ng));
  with_vec_type(dtype, [&](auto* tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    halo_pack_unpack<T><<<g, 256, 0, s>>>(m->d_indexbuf.as<int32_t>(), nidx, static_cast<T*>(vec),
                                          m->d_send.as<T>(), static_cast<const T*>(inbox),
                                          m->local_size, ng);
  });
  check_launch("halo_pack_unpack");
}

} // namespace spmvx

using namespace spmvx;

extern "C" {

int spmvx_l2gmap_create(int32_t nranks, int32_t rank, const int64_t* local_sizes,
                        const int64_t* ghosts, int64_t nghosts, const int64_t* remote_ptr,
                        const int64_t* remote_ghosts, spmvx_l2gmap* out)
{
  return guarded([&] {
    SPMVX_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    SPMVX_REQUIRE(nranks >= 1 && rank >= 0 && rank < nranks, "invalid rank / nranks");
    SPMVX_REQUIRE(local_sizes != nullptr, "local_sizes is NULL");
    SPMVX_REQUIRE(nghosts >= 0 && (nghosts == 0 || ghosts != nullptr), "invalid ghosts");
    auto m = std::make_unique<spmvx_l2gmap_s>();
    static std::atomic<uint64_t> next_id{1};
    m->id = next_id.fetch_add(1);
    m->nranks = nranks;
    m->rank = rank;
    // ranges: Allgather of local sizes + prefix sum (L2GMap.cpp:23-28)
    m->ranges.assign(nranks + 1, 0);
    for (int q = 0; q < nranks; ++q)
    {
      SPMVX_REQUIRE(local_sizes[q] >= 0, "negative local size");
      m->ranges[q + 1] = m->ranges[q] + local_sizes[q];
    }
    const int64_t r0 = m->ranges[rank], r1 = m->ranges[rank + 1];
    m->local_size = r1 - r0;
    SPMVX_REQUIRE(m->local_size < (int64_t(1) << 31), "local size too large");
    m->ghosts.assign(ghosts, ghosts + nghosts);
    if (!std::is_sorted(m->ghosts.begin(), m->ghosts.end())) // :34-35
      throw Error(SPMVX_ERR_ARG, "Ghosts must be sorted");
    // owner of each ghost (:40-54)
    std::vector<int32_t> ghost_count(nranks, 0);
    for (int64_t idx : m->ghosts)
    {
      if (idx >= r0 && idx < r1)
        throw Error(SPMVX_ERR_ARG, "Ghost index in local range"); // :44-45
      if (idx < 0 || idx >= m->ranges.back())
        throw Error(SPMVX_ERR_ARG, "Ghost index outside the global range");
      const int p = static_cast<int>(
          std::upper_bound(m->ranges.begin(), m->ranges.end(), idx) - m->ranges.begin() - 1);
      ++ghost_count[p];
    }
    // the virtual ranks' ghosts owned here (the Alltoall of :60 and the
    // Neighbor_alltoallv of :112)
    std::vector<int32_t> remote_count(nranks, 0);
    if (remote_ptr)
    {
      SPMVX_REQUIRE(remote_ptr[0] == 0, "remote_ptr[0] must be 0");
      for (int q = 0; q < nranks; ++q)
      {
        const int64_t a = remote_ptr[q], e = remote_ptr[q + 1];
        SPMVX_REQUIRE(e >= a, "remote_ptr must be non-decreasing");
        SPMVX_REQUIRE(q != rank || e == a, "a rank cannot ghost its own indices");
        SPMVX_REQUIRE(e == a || remote_ghosts != nullptr, "remote_ghosts is NULL");
        for (int64_t j = a; j < e; ++j)
        {
          SPMVX_REQUIRE(remote_ghosts[j] >= r0 && remote_ghosts[j] < r1,
                        "remote ghost not owned by this rank");
          SPMVX_REQUIRE(j == a || remote_ghosts[j] > remote_ghosts[j - 1],
                        "Ghosts must be sorted");
        }
        remote_count[q] = static_cast<int32_t>(e - a);
      }
    }
    // symmetrised neighbour list (:63-78)
    for (int q = 0; q < nranks; ++q)
    {
      if (ghost_count[q] > 0)
      {
        m->neighbours.push_back(q);
        m->send_count.push_back(ghost_count[q]);
        m->recv_count.push_back(remote_count[q]);
      }
      else if (remote_count[q] > 0)
      {
        m->neighbours.push_back(q);
        m->send_count.push_back(0);
        m->recv_count.push_back(remote_count[q]);
      }
    }
    if (m->neighbours.empty()) // :87-92
    {
      m->send_count = {0};
      m->recv_count = {0};
    }
    m->send_offset = {0};
    for (int32_t c : m->send_count)
      m->send_offset.push_back(m->send_offset.back() + c);
    m->recv_offset = {0};
    for (int32_t c : m->recv_count)
      m->recv_offset.push_back(m->recv_offset.back() + c);
    for (int q : m->neighbours)
      if (remote_ptr)
        for (int64_t j = remote_ptr[q]; j < remote_ptr[q + 1]; ++j)
          m->indexbuf.push_back(static_cast<int32_t>(remote_ghosts[j] - r0));
    for (int32_t& s : m->send_offset) // :120-121
      s += static_cast<int32_t>(m->local_size);
    // reverse_update scatter plan: group positions by target, ascending
    {
      const size_t n = m->indexbuf.size();
      std::vector<int32_t> order(n);
      for (size_t i = 0; i < n; ++i)
        order[i] = static_cast<int32_t>(i);
      std::stable_sort(order.begin(), order.end(),
                       [&](int32_t a, int32_t b) { return m->indexbuf[a] < m->indexbuf[b]; });
      m->rev_ptr.push_back(0);
      for (size_t j = 0; j < n; ++j)
      {
        const int32_t tgt = m->indexbuf[order[j]];
        if (m->rev_target.empty() || m->rev_target.back() != tgt)
        {
          if (!m->rev_target.empty())
            m->rev_ptr.push_back(static_cast<int32_t>(j));
          m->rev_target.push_back(tgt);
        }
        m->rev_pos.push_back(order[j]);
      }
      if (!m->rev_target.empty())
        m->rev_ptr.push_back(static_cast<int32_t>(n));
    }
    *out = m.release();
  });
}

int spmvx_l2gmap_destroy(spmvx_l2gmap m)
{
  return guarded([&] { delete m; });
}

int spmvx_l2gmap_info(spmvx_l2gmap m, int64_t* local_size, int64_t* num_ghosts,
                      int64_t* global_size, int64_t* global_offset, int64_t* num_indices,
                      int32_t* num_neighbours)
{
  return guarded([&] {
    SPMVX_REQUIRE(m != nullptr, "NULL map");
    if (local_size) *local_size = m->local_size;
    if (num_ghosts) *num_ghosts = static_cast<int64_t>(m->ghosts.size());
    if (global_size) *global_size = m->ranges.back();
    if (global_offset) *global_offset = m->ranges[m->rank];
    if (num_indices) *num_indices = static_cast<int64_t>(m->indexbuf.size());
    if (num_neighbours) *num_neighbours = static_cast<int32_t>(m->neighbours.size());
  });
}

int spmvx_l2gmap_plan(spmvx_l2gmap m, int32_t* neighbours, int32_t* send_count,
                      int32_t* recv_count, int32_t* send_offset, int32_t* recv_offset,
                      int32_t* indexbuf)
{
  return guarded([&] {
    SPMVX_REQUIRE(m != nullptr, "NULL map");
    auto cp = [](int32_t* dst, const std::vector<int32_t>& v) {
      if (dst && !v.empty())
        std::memcpy(dst, v.data(), 4 * v.size());
    };
    cp(neighbours, m->neighbours);
    cp(send_count, m->send_count);
    cp(recv_count, m->recv_count);
    cp(send_offset, m->send_offset);
    cp(recv_offset, m->recv_offset);
    cp(indexbuf, m->indexbuf);
  });
}

int spmvx_l2gmap_ghosts(spmvx_l2gmap m, int64_t* ghosts)
{
  return guarded([&] {
    SPMVX_REQUIRE(m != nullptr, "NULL map");
    if (ghosts && !m->ghosts.empty())
      std::memcpy(ghosts, m->ghosts.data(), 8 * m->ghosts.size());
  });
}

int spmvx_l2gmap_global_to_local(spmvx_l2gmap m, int64_t i, int32_t* local)
{
  return guarded([&] {
    SPMVX_REQUIRE(m != nullptr && local != nullptr, "NULL argument");
    const int64_t r0 = m->ranges[m->rank], r1 = m->ranges[m->rank + 1];
    if (i >= r0 && i < r1)
    {
      *local = static_cast<int32_t>(i - r0);
      return;
    }
    auto it = std::lower_bound(m->ghosts.begin(), m->ghosts.end(), i);
    if (it == m->ghosts.end() || *it != i)
      throw Error(SPMVX_ERR_ARG, "global index is neither owned nor a ghost");
    *local = static_cast<int32_t>(m->local_size + (it - m->ghosts.begin()));
  });
}

int spmvx_l2gmap_set_halo(spmvx_l2gmap m, int32_t dtype, const void* halo_host)
{
  return guarded([&] {
    SPMVX_REQUIRE(m != nullptr, "NULL map");
    if (halo_host == nullptr)
    {
      m->halo_dtype = -1;
      return;
    }
    const size_t es = vec_elem(dtype);
    const size_t ng = m->ghosts.size();
    m->d_halo.ensure(es * std::max<size_t>(ng, 1));
    if (ng)
      SPMVX_HIP(hipMemcpy(m->d_halo.ptr, halo_host, es * ng, hipMemcpyHostToDevice));
    m->halo_dtype = dtype;
  });
}

int spmvx_l2gmap_update(spmvx_l2gmap m, int32_t dtype, void* vec, void* send_host,
                        spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(m != nullptr, "NULL map");
    const size_t es = vec_elem(dtype);
    SPMVX_REQUIRE(vec != nullptr || (m->local_size == 0 && m->ghosts.empty()), "vec is NULL");
    hipStream_t s = as_stream(stream);
    l2gmap_update_async(m, dtype, vec, s);
    const size_t nidx = m->indexbuf.size();
    if (send_host)
    {
      if (nidx)
      {
        m->h_send.ensure(es * nidx);
        SPMVX_HIP(hipMemcpyAsync(m->h_send.ptr, m->d_send.ptr, es * nidx, hipMemcpyDeviceToHost, s));
      }
      SPMVX_HIP(hipStreamSynchronize(s));
      if (nidx)
        std::memcpy(send_host, m->h_send.ptr, es * nidx);
    }
  });
}

int spmvx_l2gmap_reverse_update(spmvx_l2gmap m, int32_t dtype, void* vec,
                                const void* incoming_host, void* ghost_out_host,
                                spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(m != nullptr, "NULL map");
    const size_t es = vec_elem(dtype);
    l2gmap_prepare_device(m);
    hipStream_t s = as_stream(stream);
    const size_t ng = m->ghosts.size(), nidx = m->indexbuf.size();
    SPMVX_REQUIRE(vec != nullptr || (m->local_size == 0 && ng == 0), "vec is NULL");
    // send: ghost region to the owners (L2GMap.cpp:161-164)
    if (ghost_out_host && ng)
    {
      m->h_send.ensure(es * std::max(ng, nidx));
      SPMVX_HIP(hipMemcpyAsync(m->h_send.ptr, static_cast<char*>(vec) + es * m->local_size,
                               es * ng, hipMemcpyDeviceToHost, s));
      SPMVX_HIP(hipStreamSynchronize(s));
      std::memcpy(ghost_out_host, m->h_send.ptr, es * ng);
    }
    // receive + accumulate (:168-169)
    if (incoming_host && nidx)
    {
      m->h_in.ensure(es * nidx);
      std::memcpy(m->h_in.ptr, incoming_host, es * nidx);
      SPMVX_HIP(hipMemcpyAsync(m->d_send.ptr, m->h_in.ptr, es * nidx, hipMemcpyHostToDevice, s));
      const int64_t nt = static_cast<int64_t>(m->rev_target.size());
      const int g = grid_for(nt);
      with_vec_type(dtype, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        reverse_scatter_add<T><<<g, 256, 0, s>>>(
            m->d_rev_target.as<int32_t>(), m->d_rev_ptr.as<int32_t>(), m->d_rev_pos.as<int32_t>(),
            nt, m->d_send.as<T>(), static_cast<T*>(vec));
      });
      check_launch("reverse_scatter_add");
    }
    SPMVX_HIP(hipStreamSynchronize(s));
  });
}

} // extern "C"
