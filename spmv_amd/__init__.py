"""spmv_amd -- MI355X (gfx950) implementation of libspmv's hot path.

Python front end of ``libspmv_amd.so`` (C ABI declared in
``include/spmv_amd.h``).  Mirrors the reference's operator interface:

=========================  ==============================================
reference (C++)            here
=========================  ==============================================
``spmv::L2GMap``           :class:`L2GMap` (spmv/L2GMap.h:29-97)
``spmv::Matrix<T>``        :class:`Matrix` (spmv/Matrix.h:23-81)
``Matrix::operator*``      ``A @ x`` / :meth:`Matrix.mult`
``Matrix::transpmult``     :meth:`Matrix.transpmult`
``spmv::cg``               :func:`cg` (spmv/cg.h:27-29)
=========================  ==============================================

Device vectors are torch CUDA tensors (torch is used for device memory and
streams only); every operation runs in the library's HIP kernels on the
current torch stream.  There is no CPU fallback: without the built library
or a GPU, operations raise.
"""
import ctypes
import os

import numpy as np

__all__ = ["L2GMap", "Matrix", "Batch", "cg", "cg_noalloc", "cg_f32", "ir_cg", "sddmm", "copy_f64", "read_f64", "SpmvError",
           "read_petsc_csr", "read_petsc_matrix", "read_petsc_vector",
           "F64", "F32", "BF16", "C64", "C32", "SLICED_VARIANT", "lib", "library_path"]

F64, F32, BF16, C64, C32 = 0, 1, 2, 3, 4
# kernel variant of the sliced row-pattern layout (SPMVX_SLICED_VARIANT)
SLICED_VARIANT = 12
_DTYPES = {"f64": F64, "float64": F64, "double": F64, F64: F64,
           "f32": F32, "float32": F32, "float": F32, F32: F32,
           "bf16": BF16, "bfloat16": BF16, BF16: BF16,
           "c64": C64, "complex128": C64, C64: C64,
           "c32": C32, "complex64": C32, C32: C32}
# numpy element type of each vector dtype
_NP_VEC = {F64: np.float64, F32: np.float32, C64: np.complex128, C32: np.complex64}


def _vec_dtype_of(t):
    """spmvx dtype of a torch vector."""
    import torch
    m = {torch.float64: F64, torch.float32: F32, torch.complex128: C64, torch.complex64: C32}
    if t.dtype not in m:
        raise SpmvError(f"unsupported vector dtype {t.dtype}")
    return m[t.dtype]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libspmv_amd.so")
_LIB = None


class SpmvError(RuntimeError):
    """Raised when a library call fails (the reference throws runtime_error)."""


class IRStats(ctypes.Structure):
    _fields_ = [("outer_its", ctypes.c_int32), ("inner_its", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("outer_hist", ctypes.c_double * 64),
                ("inner_hist", ctypes.c_int32 * 64)]

    def as_dict(self):
        n = self.outer_its
        return dict(outer_its=n, inner_its=self.inner_its, rel_residual=self.rel_residual,
                    outer_hist=list(self.outer_hist[: n + 1]), inner_hist=list(self.inner_hist[:n]))


def lib():
    """Loads libspmv_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(library_path):
        raise ImportError(f"{library_path} is missing: build it with "
                          "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(library_path)
    P = ctypes.c_void_p
    i32, i64, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "spmvx_last_error": ([], ctypes.c_char_p),
        "spmvx_abi_version": ([], ctypes.c_int),
        "spmvx_l2gmap_create": ([i32, i32, P, P, i64, P, P, P], ctypes.c_int),
        "spmvx_l2gmap_destroy": ([P], ctypes.c_int),
        "spmvx_l2gmap_info": ([P, P, P, P, P, P, P], ctypes.c_int),
        "spmvx_l2gmap_plan": ([P, P, P, P, P, P, P], ctypes.c_int),
        "spmvx_l2gmap_ghosts": ([P, P], ctypes.c_int),
        "spmvx_l2gmap_global_to_local": ([P, i64, P], ctypes.c_int),
        "spmvx_l2gmap_set_halo": ([P, i32, P], ctypes.c_int),
        "spmvx_l2gmap_update": ([P, i32, P, P, P], ctypes.c_int),
        "spmvx_l2gmap_reverse_update": ([P, i32, P, P, P, P], ctypes.c_int),
        "spmvx_matrix_create": ([i32, i64, i64, i64, P, P, P, P], ctypes.c_int),
        "spmvx_matrix_create_like": ([P, i32, P], ctypes.c_int),
        "spmvx_matrix_destroy": ([P], ctypes.c_int),
        "spmvx_matrix_info": ([P, P, P, P, P, P], ctypes.c_int),
        "spmvx_matrix_arrays": ([P, P, P, P], ctypes.c_int),
        "spmvx_matrix_set_values": ([P, P, P], ctypes.c_int),
        "spmvx_matrix_set_kernel": ([P, i32], ctypes.c_int),
        "spmvx_matrix_set_long_rows": ([P, i32], ctypes.c_int),
        "spmvx_num_kernels": ([], ctypes.c_int),
        "spmvx_matrix_sliced_info": ([P, P, P, P], ctypes.c_int),
        "spmvx_matrix_get_kernel": ([P, P, ctypes.c_char_p, i32], ctypes.c_int),
        "spmvx_spmv": ([P, P, P, P], ctypes.c_int),
        "spmvx_spmv_t": ([P, P, P, P], ctypes.c_int),
        "spmvx_spmv_timed": ([P, P, P, P, P, P], ctypes.c_int),
        "spmvx_cg": ([P, P, P, P, i32, f64, P, P, P], ctypes.c_int),
        "spmvx_cg_noalloc": ([P, P, P, P, i32, f64, P, P, P], ctypes.c_int),
        "spmvx_cg_f32": ([P, P, P, i32, f64, P, P], ctypes.c_int),
        "spmvx_ir_cg": ([P, P, P, P, i32, f64, i32, f64, i32, P, P], ctypes.c_int),
        "spmvx_sddmm": ([P, f64, P, P, P, P], ctypes.c_int),
        "spmvx_copy_f64": ([P, P, i64, P], ctypes.c_int),
        "spmvx_read_f64": ([P, i64, P, P], ctypes.c_int),
        "spmvx_petsc_csr_read": ([ctypes.c_char_p, i32, i32, P], ctypes.c_int),
        "spmvx_petsc_csr_destroy": ([P], ctypes.c_int),
        "spmvx_petsc_csr_info": ([P, P, P, P, P, P, P], ctypes.c_int),
        "spmvx_petsc_csr_arrays": ([P, P, P, P, P], ctypes.c_int),
        "spmvx_petsc_csr_remote": ([P, P, P], ctypes.c_int),
        "spmvx_petsc_read_matrix": ([ctypes.c_char_p, i32, i32, i32, P, P, P], ctypes.c_int),
        "spmvx_petsc_read_vector": ([ctypes.c_char_p, i32, i32, P, i64, P, P], ctypes.c_int),
        "spmvx_batch_create": ([i32, P, P, P, P, i32, P], ctypes.c_int),
        "spmvx_batch_destroy": ([P], ctypes.c_int),
        "spmvx_batch_set_config": ([P, i32, i32, i32], ctypes.c_int),
        "spmvx_batch_solve": ([P, P, P, i32, f64, P, P], ctypes.c_int),
        "spmvx_batch_info": ([P, P, P, P, P, P, P, P, P, P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def _check(status):
    if status != 0:
        msg = lib().spmvx_last_error().decode()
        raise SpmvError(msg or f"spmv_amd error {status}")


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _stream(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def _dptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise SpmvError("device vector expected (a CUDA tensor)")
    if not t.is_contiguous():
        raise SpmvError("contiguous vector expected")
    return ctypes.c_void_p(t.data_ptr())


def _vec_torch_dtype(dtype):
    import torch
    return {F64: torch.float64, C64: torch.complex128, C32: torch.complex64}.get(dtype, torch.float32)


# =========================================================================
class L2GMap:
    """Local-to-global map with ghost entries (spmv/L2GMap.h:29-97).

    ``L2GMap(local_size, ghosts)`` is the one-rank map of the reference's
    constructor ``L2GMap(comm, local_size, ghosts)``.  A rank of a larger
    decomposition is described with ``local_sizes`` (all ranks' sizes),
    ``rank`` and ``remote_ghosts`` (for each rank q, the sorted global
    indices of q's ghosts owned here); the exchange plan then equals what the
    reference computes with MPI.  The neighbours' data for ``update`` comes
    from :meth:`set_halo` (zeros by default).
    """

    def __init__(self, local_size, ghosts=(), *, local_sizes=None, rank=0, remote_ghosts=None):
        if local_sizes is None:
            local_sizes = [int(local_size)]
            rank = 0
        local_sizes = np.ascontiguousarray(local_sizes, np.int64)
        if int(local_sizes[rank]) != int(local_size):
            raise SpmvError("local_size differs from local_sizes[rank]")
        ghosts = np.ascontiguousarray(ghosts, np.int64)
        nranks = len(local_sizes)
        rptr = ridx = None
        if remote_ghosts is not None:
            lens = [len(remote_ghosts.get(q, ())) if isinstance(remote_ghosts, dict)
                    else len(remote_ghosts[q]) for q in range(nranks)]
            rptr = np.zeros(nranks + 1, np.int64)
            rptr[1:] = np.cumsum(lens)
            seq = [remote_ghosts.get(q, ()) if isinstance(remote_ghosts, dict) else remote_ghosts[q]
                   for q in range(nranks)]
            ridx = np.ascontiguousarray(np.concatenate([np.asarray(s, np.int64) for s in seq])
                                        if rptr[-1] else np.zeros(1, np.int64))
        h = ctypes.c_void_p()
        _check(lib().spmvx_l2gmap_create(nranks, rank, _np_ptr(local_sizes), _np_ptr(ghosts),
                                         len(ghosts), _np_ptr(rptr), _np_ptr(ridx),
                                         ctypes.byref(h)))
        self._h = h
        self._keep = (rptr, ridx)

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.spmvx_l2gmap_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def _info(self):
        v = [ctypes.c_int64() for _ in range(5)] + [ctypes.c_int32()]
        _check(lib().spmvx_l2gmap_info(self._h, *[ctypes.byref(x) for x in v]))
        return [x.value for x in v]

    def local_size(self):
        return self._info()[0]

    def num_ghosts(self):
        return self._info()[1]

    def global_size(self):
        return self._info()[2]

    def global_offset(self):
        return self._info()[3]

    def num_indices(self):
        return self._info()[4]

    def ghosts(self):
        out = np.zeros(max(self.num_ghosts(), 1), np.int64)
        _check(lib().spmvx_l2gmap_ghosts(self._h, _np_ptr(out)))
        return out[: self.num_ghosts()]

    def global_to_local(self, i):
        out = ctypes.c_int32()
        _check(lib().spmvx_l2gmap_global_to_local(self._h, int(i), ctypes.byref(out)))
        return out.value

    def plan(self):
        ls, ng, gs, go, ni, nn = self._info()
        m = max(nn, 1)
        arrs = dict(neighbours=np.zeros(max(nn, 1), np.int32), send_count=np.zeros(m, np.int32),
                    recv_count=np.zeros(m, np.int32), send_offset=np.zeros(m + 1, np.int32),
                    recv_offset=np.zeros(m + 1, np.int32), indexbuf=np.zeros(max(ni, 1), np.int32))
        _check(lib().spmvx_l2gmap_plan(self._h, *[_np_ptr(a) for a in arrs.values()]))
        arrs["neighbours"] = arrs["neighbours"][:nn]
        arrs["indexbuf"] = arrs["indexbuf"][:ni]
        return {k: v.tolist() for k, v in arrs.items()}

    def set_halo(self, halo, dtype=F64):
        """Host values that the neighbours deliver into the ghost region."""
        if halo is None:
            _check(lib().spmvx_l2gmap_set_halo(self._h, F64, None))
            return
        halo = np.ascontiguousarray(halo, _NP_VEC[_DTYPES[dtype]])
        if len(halo) != self.num_ghosts():
            raise SpmvError("halo length differs from num_ghosts")
        _check(lib().spmvx_l2gmap_set_halo(self._h, dtype, _np_ptr(halo)))

    def update(self, vec, send_out=False, stream=None):
        """L2GMap::update on a device vector (in place).  Returns the packed
        send buffer (numpy) when send_out is True."""
        dt = _vec_dtype_of(vec)
        if vec.numel() != self.local_size() + self.num_ghosts():
            raise SpmvError("vector length differs from local_size + num_ghosts")
        buf = None
        if send_out:
            buf = np.zeros(max(self.num_indices(), 1), _NP_VEC[dt])
        _check(lib().spmvx_l2gmap_update(self._h, dt, _dptr(vec), _np_ptr(buf), _stream(stream)))
        return None if buf is None else buf[: self.num_indices()]

    def reverse_update(self, vec, incoming=None, stream=None):
        """L2GMap::reverse_update: returns the ghost values sent to the owners
        and scatter-adds `incoming` (host, |indexbuf| values) into vec."""
        dt = _vec_dtype_of(vec)
        if vec.numel() != self.local_size() + self.num_ghosts():
            raise SpmvError("vector length differs from local_size + num_ghosts")
        npdt = _NP_VEC[dt]
        out = np.zeros(max(self.num_ghosts(), 1), npdt)
        inc = None if incoming is None else np.ascontiguousarray(incoming, npdt)
        if inc is not None and len(inc) != self.num_indices():
            raise SpmvError("incoming length differs from the index buffer")
        _check(lib().spmvx_l2gmap_reverse_update(self._h, dt, _dptr(vec), _np_ptr(inc),
                                                 _np_ptr(out), _stream(stream)))
        return out[: self.num_ghosts()]


# =========================================================================
class Matrix:
    """Distributed CSR matrix block (spmv/Matrix.h:23-81) on the device.

    ``Matrix(rowptr, colidx, values, ncols, col_map, row_map, dtype)`` copies
    the compressed row-major CSR in (as ``Matrix<T>::Matrix`` copies the
    Eigen matrix, Matrix.cpp:17) and runs the row-binning analysis.
    """

    def __init__(self, rowptr, colidx, values, ncols=None, col_map=None, row_map=None,
                 dtype="f64", _handle=None):
        if _handle is not None:
            self._h = _handle
        else:
            rowptr = np.ascontiguousarray(rowptr, np.int32)
            colidx = np.ascontiguousarray(colidx, np.int32)
            if _DTYPES[dtype] in (C64, C32):  # (re, im) pairs of doubles
                values = np.ascontiguousarray(values, np.complex128).view(np.float64)
            else:
                values = np.ascontiguousarray(values, np.float64)
            nrows = len(rowptr) - 1
            if ncols is None:
                ncols = col_map.local_size() + col_map.num_ghosts() if col_map else nrows
            h = ctypes.c_void_p()
            _check(lib().spmvx_matrix_create(_DTYPES[dtype], nrows, int(ncols), len(colidx),
                                             _np_ptr(rowptr), _np_ptr(colidx), _np_ptr(values),
                                             ctypes.byref(h)))
            self._h = h
        self._col_map = col_map
        self._row_map = row_map
        self._base = None
        dt, nr, nc, nz, nb = [ctypes.c_int32()] + [ctypes.c_int64() for _ in range(4)]
        _check(lib().spmvx_matrix_info(self._h, *[ctypes.byref(x) for x in (dt, nr, nc, nz, nb)]))
        self.dtype, self.nrows, self.ncols, self.nnz, self.nblocks = (
            dt.value, nr.value, nc.value, nz.value, nb.value)

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.spmvx_matrix_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def like(self, dtype):
        """Same pattern (shared on the device), values converted to dtype."""
        h = ctypes.c_void_p()
        _check(lib().spmvx_matrix_create_like(self._h, _DTYPES[dtype], ctypes.byref(h)))
        m = Matrix(None, None, None, _handle=h, col_map=self._col_map, row_map=self._row_map)
        m._base = self  # keeps the shared pattern's owner alive
        return m

    def row_map(self):
        return self._row_map

    def col_map(self):
        return self._col_map

    def vector_dtype(self):
        return _vec_torch_dtype(self.dtype)

    def mult(self, x, out=None, stream=None):
        """y = A x (Matrix<T>::operator*, Matrix.cpp:199-204)."""
        import torch
        if x.numel() != self.ncols:
            raise SpmvError("x has the wrong length")
        if x.dtype != self.vector_dtype():
            raise SpmvError("x has the wrong dtype")
        if out is None:
            out = torch.empty(self.nrows, dtype=self.vector_dtype(), device=x.device)
        elif out.numel() != self.nrows or out.dtype != self.vector_dtype():
            raise SpmvError("out has the wrong length or dtype")
        _check(lib().spmvx_spmv(self._h, _dptr(x), _dptr(out), _stream(stream)))
        return out

    __matmul__ = mult

    def mult_timed(self, x, out, start, stop, stream=None):
        """mult(x, out) with two torch.cuda.Event(enable_timing=True) that
        record the product kernel's own start and end: start.elapsed_time(stop)
        is then the kernel's execution time without its dispatch."""
        import torch
        if x.numel() != self.ncols or x.dtype != self.vector_dtype():
            raise SpmvError("x has the wrong length or dtype")
        if out.numel() != self.nrows or out.dtype != self.vector_dtype():
            raise SpmvError("out has the wrong length or dtype")
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        for ev in (start, stop):
            if not ev.cuda_event:  # torch creates its events at the first record
                ev.record(s)
        _check(lib().spmvx_spmv_timed(self._h, _dptr(x), _dptr(out), _stream(s),
                                      ctypes.c_void_p(start.cuda_event),
                                      ctypes.c_void_p(stop.cuda_event)))
        return out

    def transpmult(self, x, out=None, stream=None):
        """y = A^T x (Matrix<T>::transpmult, Matrix.cpp:206-211)."""
        import torch
        if x.numel() != self.nrows:
            raise SpmvError("x has the wrong length")
        if x.dtype != self.vector_dtype():
            raise SpmvError("x has the wrong dtype")
        if out is None:
            out = torch.empty(self.ncols, dtype=self.vector_dtype(), device=x.device)
        elif out.numel() != self.ncols or out.dtype != self.vector_dtype():
            raise SpmvError("out has the wrong length or dtype")
        _check(lib().spmvx_spmv_t(self._h, _dptr(x), _dptr(out), _stream(stream)))
        return out

    def set_kernel(self, variant):
        """SpMV kernel variant (-1 = default); results are identical.
        ``SLICED_VARIANT`` selects the sliced row-pattern layout and raises
        SpmvError with the reason when the matrix does not fit it."""
        _check(lib().spmvx_matrix_set_kernel(self._h, int(variant)))

    def set_long_rows(self, parallel):
        """Rows longer than a kernel's staging block: summed in stored order by
        one lane (False, the default: bit-identical to the sequential product)
        or by the whole workgroup in a fixed tree (True: deterministic, rounded
        differently, much faster for very long rows).  Real values only."""
        _check(lib().spmvx_matrix_set_long_rows(self._h, 1 if parallel else 0))

    def kernel(self):
        """(variant, description) of the SpMV kernel this matrix uses."""
        v = ctypes.c_int32()
        buf = ctypes.create_string_buffer(160)
        _check(lib().spmvx_matrix_get_kernel(self._h, ctypes.byref(v), buf, len(buf)))
        return v.value, buf.value.decode()

    def kernel_name(self):
        return "variant %d: %s" % self.kernel()

    def sliced_info(self):
        """Row-pattern analysis for the sliced layout: dict(npatterns, wmax,
        slots); raises SpmvError with the reason when the matrix does not fit."""
        n, w, sl = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _check(lib().spmvx_matrix_sliced_info(self._h, ctypes.byref(n), ctypes.byref(w),
                                              ctypes.byref(sl)))
        return {"npatterns": n.value, "wmax": w.value, "slots": sl.value}

    def set_values(self, values_f64, stream=None):
        _check(lib().spmvx_matrix_set_values(self._h, _dptr(values_f64), _stream(stream)))

    def arrays(self):
        """Device pointers (rowptr, colidx, values)."""
        p = [ctypes.c_void_p() for _ in range(3)]
        _check(lib().spmvx_matrix_arrays(self._h, *[ctypes.byref(x) for x in p]))
        return [x.value for x in p]


# =========================================================================
def cg(A, b, max_its, rtol, history=False, stream=None):
    """spmv::cg (spmv/cg.cpp:11-83): returns (x, k) or (x, k, hist) where
    hist[j] = ||r_j||^2 and x has the column-map length (ghosts zero)."""
    import torch
    rm = A.row_map()
    if rm is not None and rm.num_ghosts() > 0:
        raise SpmvError("spmv::cg - Error: A.row_map() has ghost entries")
    if b.numel() != A.nrows:
        raise SpmvError("spmv::cg - Error: b.rows() != A.rows()")
    if b.dtype != torch.float64:
        raise SpmvError("b must be float64")
    x = torch.empty(A.ncols, dtype=torch.float64, device=b.device)
    its = ctypes.c_int32()
    hist = np.zeros(max_its + 1, np.float64) if history else None
    cm = A.col_map()
    _check(lib().spmvx_cg(A.handle, cm.handle if cm is not None else None, _dptr(b), _dptr(x),
                          int(max_its), float(rtol), ctypes.byref(its), _np_ptr(hist),
                          _stream(stream)))
    if history:
        return x, its.value, hist[: its.value + 1]
    return x, its.value


def cg_noalloc(A, b, max_its, rtol, x=None, history=False, stream=None):
    """spmvx_cg_noalloc: spmv::cg with every buffer of the solve in the
    library's fixed device scratch (no allocation by the library; x may be
    given so that the caller allocates nothing either).  Returns (x, k) or
    (x, k, hist) like cg()."""
    import torch
    if b.numel() != A.nrows:
        raise SpmvError("spmv::cg - Error: b.rows() != A.rows()")
    if b.dtype != torch.float64:
        raise SpmvError("b must be float64")
    if x is None:
        x = torch.empty(A.ncols, dtype=torch.float64, device=b.device)
    elif x.numel() != A.ncols or x.dtype != torch.float64:
        raise SpmvError("x must hold A.ncols float64 values")
    its = ctypes.c_int32()
    hist = np.zeros(max_its + 1, np.float64) if history else None
    cm = A.col_map()
    _check(lib().spmvx_cg_noalloc(A.handle, cm.handle if cm is not None else None, _dptr(b),
                                  _dptr(x), int(max_its), float(rtol), ctypes.byref(its),
                                  _np_ptr(hist), _stream(stream)))
    if history:
        return x, its.value, hist[: its.value + 1]
    return x, its.value


def cg_f32(A, b, max_its, rtol, stream=None):
    """FP32 CG on an F32/BF16 matrix: returns (x, k)."""
    import torch
    if b.numel() != A.nrows or b.dtype != torch.float32:
        raise SpmvError("b must be float32 with A.nrows entries")
    x = torch.empty(A.ncols, dtype=torch.float32, device=b.device)
    its = ctypes.c_int32()
    _check(lib().spmvx_cg_f32(A.handle, _dptr(b), _dptr(x), int(max_its), float(rtol),
                              ctypes.byref(its), _stream(stream)))
    return x, its.value


def ir_cg(A64, Alow, b, rtol=1e-10, inner_rtol=1e-4, max_outer=32, inner_max_its=100000,
          stream=None, adaptive_inner=False):
    """Mixed-precision iterative refinement: returns (x, stats dict).
    adaptive_inner: each inner solve stops at max(inner_rtol, rtol / (2 rel))
    with rel the current outer relative residual (SPMVX_IR_ADAPTIVE_INNER)."""
    import torch
    if b.numel() != A64.nrows or b.dtype != torch.float64:
        raise SpmvError("b must be float64 with A.nrows entries")
    x = torch.empty(A64.ncols, dtype=torch.float64, device=b.device)
    st = IRStats()
    _check(lib().spmvx_ir_cg(A64.handle, Alow.handle, _dptr(b), _dptr(x), int(max_outer),
                             float(rtol), int(inner_max_its), float(inner_rtol),
                             1 if adaptive_inner else 0, ctypes.byref(st), _stream(stream)))
    return x, st.as_dict()


def sddmm(A, u, v, alpha=1.0, out=None, stream=None):
    """out[k] = alpha * u[row(k)] * v[col(k)] on A's pattern."""
    import torch
    if u.numel() != A.nrows or v.numel() != A.ncols:
        raise SpmvError("u must have A.nrows and v A.ncols entries")
    if out is None:
        out = torch.empty(A.nnz, dtype=torch.float64, device=u.device)
    elif out.numel() != A.nnz:
        raise SpmvError("out must have A.nnz entries")
    _check(lib().spmvx_sddmm(A.handle, float(alpha), _dptr(u), _dptr(v), _dptr(out),
                             _stream(stream)))
    return out


def copy_f64(dst, src, stream=None):
    if dst.numel() < src.numel():
        raise SpmvError("dst is shorter than src")
    _check(lib().spmvx_copy_f64(_dptr(dst), _dptr(src), src.numel(), _stream(stream)))


def read_f64(src, out, stream=None):
    """out[0] = sum of the float64 vector src, read once with streaming
    (non-temporal) loads: the streaming-read ceiling of the GPU."""
    import torch
    if src.dtype != torch.float64 or out.dtype != torch.float64:
        raise SpmvError("read_f64 takes float64 tensors")
    if out.numel() < 1:
        raise SpmvError("out must hold one value")
    _check(lib().spmvx_read_f64(_dptr(src), src.numel(), _dptr(out), _stream(stream)))


# =========================================================================
class Batch:
    """Many independent SPD systems solved by one persistent CG launch.

    ``systems`` is a list of host CSRs (objects with rowptr, colidx, val,
    nrows); they are stacked into one block-diagonal matrix.  Vectors are
    torch float32 tensors of shape (total_rows, nrhs).
    """

    def __init__(self, systems, nrhs=4):
        sys_row = np.zeros(len(systems) + 1, np.int64)
        sys_row[1:] = np.cumsum([A.nrows for A in systems])
        nnz = np.zeros(len(systems) + 1, np.int64)
        nnz[1:] = np.cumsum([int(A.rowptr[-1]) for A in systems])
        if sys_row[-1] >= 2 ** 31 or nnz[-1] >= 2 ** 31:
            raise SpmvError("batch too large for 32-bit indices")
        rowptr = np.concatenate([np.asarray(A.rowptr[:-1], np.int64) + nnz[s]
                                 for s, A in enumerate(systems)] + [nnz[-1:]]).astype(np.int32)
        colidx = np.concatenate([np.asarray(A.colidx, np.int64) + sys_row[s]
                                 for s, A in enumerate(systems)]).astype(np.int32)
        vals = np.concatenate([np.asarray(A.val, np.float32) for A in systems])
        self.sys_row = sys_row.astype(np.int32)
        self.nsys, self.nrhs = len(systems), int(nrhs)
        self.nrows, self.nnz = int(sys_row[-1]), int(nnz[-1])
        h = ctypes.c_void_p()
        _check(lib().spmvx_batch_create(self.nsys, _np_ptr(self.sys_row), _np_ptr(rowptr),
                                        _np_ptr(colidx), _np_ptr(vals), self.nrhs, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.spmvx_batch_destroy(self._h)
            self._h = None

    def set_config(self, block, lanes_per_row, rows_per_member=None):
        if rows_per_member is None:
            rows_per_member = self.config()["rows_per_member"]
        _check(lib().spmvx_batch_set_config(self._h, int(block), int(lanes_per_row),
                                            int(rows_per_member)))

    def config(self):
        v = [ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32(),
             ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(),
             ctypes.c_int32()]
        _check(lib().spmvx_batch_info(self._h, *[ctypes.byref(x) for x in v]))
        return dict(block=v[4].value, lanes_per_row=v[5].value, rows_per_member=v[6].value,
                    grid=v[7].value, slots=v[8].value)

    def solve(self, b, x=None, max_its=10000, rtol=1e-6, iterations=True, stream=None):
        """Returns x (and the (nsys, nrhs) iteration counts when iterations)."""
        import torch
        if b.dtype != torch.float32 or b.shape != (self.nrows, self.nrhs) or not b.is_contiguous():
            raise SpmvError("b must be a contiguous float32 tensor of shape (rows, nrhs)")
        if x is None:
            x = torch.empty_like(b)
        elif x.shape != b.shape or x.dtype != b.dtype:
            raise SpmvError("x must match b")
        its = np.zeros(self.nsys * self.nrhs, np.int32) if iterations else None
        _check(lib().spmvx_batch_solve(self._h, _dptr(b), _dptr(x), int(max_its), float(rtol),
                                       _np_ptr(its), _stream(stream)))
        if iterations:
            return x, its.reshape(self.nsys, self.nrhs)
        return x


# ---------------------------------------------------------------- PETSc input
def _host_array(ptr, n, ctype, dtype):
    if n == 0 or not ptr.value:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (n,)).astype(dtype)


def read_petsc_csr(path, nranks=1, rank=0):
    """Host-side parse of a PETSc binary matrix file for rank `rank` of an
    `nranks`-way row split (spmv/read_petsc.cpp:32-191): dict with the local
    CSR (rowptr, colidx in local column numbering, values), ghosts (global,
    ascending), sizes, and remote (for each rank q, the columns of this rank
    that q ghosts)."""
    h = ctypes.c_void_p()
    _check(lib().spmvx_petsc_csr_read(os.fsencode(path), int(nranks), int(rank), ctypes.byref(h)))
    try:
        v = [ctypes.c_int64() for _ in range(6)]
        _check(lib().spmvx_petsc_csr_info(h, *[ctypes.byref(x) for x in v]))
        nrows, ncols, nrl, ncl, ngh, nnz = [x.value for x in v]
        p = [ctypes.c_void_p() for _ in range(4)]
        _check(lib().spmvx_petsc_csr_arrays(h, *[ctypes.byref(x) for x in p]))
        rp = [ctypes.c_void_p() for _ in range(2)]
        _check(lib().spmvx_petsc_csr_remote(h, *[ctypes.byref(x) for x in rp]))
        rptr = _host_array(rp[0], nranks + 1, ctypes.c_int64, np.int64)
        rg = _host_array(rp[1], int(rptr[-1]), ctypes.c_int64, np.int64)
        return {"nrows": nrows, "ncols": ncols, "nrows_local": nrl, "ncols_local": ncl,
                "rowptr": _host_array(p[0], nrl + 1, ctypes.c_int32, np.int64),
                "colidx": _host_array(p[1], nnz, ctypes.c_int32, np.int64),
                "values": _host_array(p[2], nnz, ctypes.c_double, np.float64),
                "ghosts": _host_array(p[3], ngh, ctypes.c_int64, np.int64),
                "remote": [rg[rptr[q]:rptr[q + 1]] for q in range(nranks)]}
    finally:
        lib().spmvx_petsc_csr_destroy(h)


def read_petsc_matrix(path, nranks=1, rank=0, dtype="f64"):
    """spmv::read_petsc_binary (read_petsc.cpp:32-191): the Matrix of rank
    `rank` with its column map (the exchange plan of an `nranks`-way row
    split) and row map attached."""
    c = read_petsc_csr(path, nranks, rank)
    cr = _owner_ranges(nranks, c["ncols"])
    rr = _owner_ranges(nranks, c["nrows"])
    col_map = L2GMap(c["ncols_local"], c["ghosts"], local_sizes=np.diff(cr), rank=rank,
                     remote_ghosts=c["remote"])
    row_map = L2GMap(c["nrows_local"], (), local_sizes=np.diff(rr), rank=rank)
    return Matrix(c["rowptr"], c["colidx"], c["values"],
                  ncols=c["ncols_local"] + len(c["ghosts"]), col_map=col_map, row_map=row_map,
                  dtype=dtype)


def read_petsc_vector(path, nranks=1, rank=0):
    """spmv::read_petsc_binary_vector (read_petsc.cpp:193-260): this rank's
    part as a host array."""
    nl, ng = ctypes.c_int64(), ctypes.c_int64()
    _check(lib().spmvx_petsc_read_vector(os.fsencode(path), int(nranks), int(rank), None, 0,
                                         ctypes.byref(nl), ctypes.byref(ng)))
    out = np.zeros(max(nl.value, 1), np.float64)
    _check(lib().spmvx_petsc_read_vector(os.fsencode(path), int(nranks), int(rank),
                                         _np_ptr(out), nl.value, None, None))
    return out[:nl.value]


def _owner_ranges(size, n):
    q, r = divmod(int(n), int(size))
    return np.array([k * (q + 1) if k < r else k * q + r for k in range(size + 1)], np.int64)
