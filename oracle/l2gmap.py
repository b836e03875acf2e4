"""L2GMap plan restated over a virtual communicator.  TEST INFRASTRUCTURE ONLY.

Restates the constructor of ``spmv::L2GMap`` (spmv/L2GMap.cpp:15-122) for
every rank of a decomposition at once, so that the collectives it performs
(``MPI_Allgather`` :25, ``MPI_Alltoall`` :60, ``MPI_Neighbor_alltoall`` :97,
``MPI_Neighbor_alltoallv`` :112) become plain list operations, plus the data
movement of ``update`` (:126-150) and ``reverse_update`` (:152-170).

Used by the tests as the checker of the library's single-rank plan builder;
itself checked against real ``mpirun`` runs of ``oracle/l2gmap_mpi.c``
(fixtures in tests/golden/l2gmap/).  Pure-Python loops: small cases only.
"""
from bisect import bisect_right


def owner_ranges(size, N):
    """demos/CreateA.cpp:12-29 and spmv/read_petsc.cpp:13-30."""
    n, r = divmod(N, size)
    return [rank * (n + 1) if rank < r else rank * n + r for rank in range(size + 1)]


def build_plans(local_sizes, ghosts_per_rank):
    """Plan of every rank; ``ghosts_per_rank[q]`` is rank q's sorted ghost list."""
    size = len(local_sizes)
    ranges = [0]
    for ls in local_sizes:  # Allgather + prefix sum, L2GMap.cpp:23-28
        ranges.append(ranges[-1] + ls)

    ghost_count = []
    ghost_local = []
    for rank in range(size):
        r0, r1 = ranges[rank], ranges[rank + 1]
        g = ghosts_per_rank[rank]
        if list(g) != sorted(g):
            raise RuntimeError("Ghosts must be sorted")  # :34-35
        cnt = [0] * size
        loc = []
        for idx in g:
            if r0 <= idx < r1:
                raise RuntimeError("Ghost index in local range")  # :44-45
            p = bisect_right(ranges, idx) - 1  # upper_bound - 1, :48-50
            if p < 0 or p >= size:
                raise RuntimeError("Ghost index outside the global range")
            cnt[p] += 1
            loc.append(idx - ranges[p])  # :53
        ghost_count.append(cnt)
        ghost_local.append(loc)

    plans = []
    for rank in range(size):
        remote_count = [ghost_count[q][rank] for q in range(size)]  # Alltoall :60
        neighbours, send_count = [], []
        for i in range(size):  # :63-78
            c, rc = ghost_count[rank][i], remote_count[i]
            if c > 0:
                neighbours.append(i)
                send_count.append(c)
            elif rc > 0:
                neighbours.append(i)
                send_count.append(0)
        # Neighbor_alltoall of send counts (:97): from neighbour q we receive
        # the count q sends to us = number of q's ghosts that we own.
        recv_count = [ghost_count[q][rank] for q in neighbours]
        if not neighbours:
            send_count, recv_count = [0], [0]  # :87-92
        send_offset = [0]
        for c in send_count:
            send_offset.append(send_offset[-1] + c)
        recv_offset = [0]
        for c in recv_count:
            recv_offset.append(recv_offset[-1] + c)
        # Neighbor_alltoallv of ghost_local (:112): from neighbour q we receive
        # q's ghost_local entries that lie in our range, in q's ghost order.
        indexbuf = []
        for q in neighbours:
            r0, r1 = ranges[rank], ranges[rank + 1]
            for gidx, lidx in zip(ghosts_per_rank[q], ghost_local[q]):
                if r0 <= gidx < r1:
                    indexbuf.append(lidx)
        send_offset = [s + local_sizes[rank] for s in send_offset]  # :120-121
        plans.append(
            dict(
                ranges=ranges,
                local_size=local_sizes[rank],
                ghosts=list(ghosts_per_rank[rank]),
                neighbours=neighbours,
                send_count=send_count,
                recv_count=recv_count,
                send_offset=send_offset,
                recv_offset=recv_offset,
                indexbuf=indexbuf,
            )
        )
    return plans


def update_all(plans, vecs):
    """``update`` on every rank at once (L2GMap.cpp:126-150), in place."""
    size = len(plans)
    sends = []
    for rank in range(size):
        pl = plans[rank]
        databuf = [vecs[rank][i] for i in pl["indexbuf"]]  # :137-139
        sends.append(databuf)
    for rank in range(size):
        pl = plans[rank]
        # Reversed roles (:141-147): we receive send_count[n] values from
        # neighbour n, landing at send_offset[n]; neighbour n sends its slice
        # recv_offset[m]..+recv_count[m] for us (m = our index in its list).
        for n, q in enumerate(pl["neighbours"]):
            plq = plans[q]
            m = plq["neighbours"].index(rank)
            chunk = sends[q][plq["recv_offset"][m]: plq["recv_offset"][m] + plq["recv_count"][m]]
            assert len(chunk) == pl["send_count"][n]
            off = pl["send_offset"][n]
            vecs[rank][off: off + len(chunk)] = chunk


def reverse_update_all(plans, vecs):
    """``reverse_update`` on every rank at once (L2GMap.cpp:152-170), in place."""
    size = len(plans)
    incoming = []
    for rank in range(size):
        pl = plans[rank]
        buf = []
        for m, q in enumerate(pl["neighbours"]):
            plq = plans[q]
            n = plq["neighbours"].index(rank)
            off = plq["send_offset"][n]
            buf.extend(vecs[q][off: off + plq["send_count"][n]])
        incoming.append(buf)
    for rank in range(size):
        pl = plans[rank]
        for i, idx in enumerate(pl["indexbuf"]):  # :168-169
            vecs[rank][idx] += incoming[rank][i]
