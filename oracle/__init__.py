"""CPU oracle for the libspmv hot path -- TEST INFRASTRUCTURE ONLY.

ctypes/numpy front end of ``oracle/liboracle.so`` (a plain-C restatement of
``Matrix<double>::operator*`` spmv/Matrix.cpp:199-204, ``L2GMap::update`` /
``reverse_update`` spmv/L2GMap.cpp:126-170 and ``spmv::cg`` spmv/cg.cpp:11-83)
and of the L2GMap plan restatement in ``oracle/l2gmap.py``.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this package, as the checker or the timed CPU reference path.
The GPU library never calls it.

Parity status per function (how each is pinned):
  spmv_f64        scipy csr_matvec (same sequential row sum, bit-exact), dense
                  brute force, closed-form eigenvectors of the Dirichlet
                  Laplacians, A*1 closed forms               -> pinned
  spmv_t_f64      dense A^T x brute force                     -> pinned
  l2g update /    real mpirun runs of oracle/l2gmap_mpi.c     -> pinned
  reverse_update  (tests/golden/l2gmap)
  cg_f64          1-iteration convergence on an eigenvector, <= m iterations
                  for m eigen-components, scipy spsolve solution, true
                  residual                                    -> pinned (to
                  tolerance: the reference's Eigen SIMD reduction order is not
                  reproducible without Eigen)
  cg_f32          same checks in single precision             -> pinned
  generators      nnz closed forms, scipy kron Laplacians     -> pinned
  adjoint         dense numpy adjoint on small grids, central differences
                                                              -> pinned
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def build():
    """Compile liboracle.so with gcc (called by __graft_entry__.build())."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_spmv_f64.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_spmv_f64_omp.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p, _f64p, ctypes.c_int]
        L.or_spmv_t_f64.argtypes = [_i64, _i64, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_spmv_f32.argtypes = [_i64, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.or_spmv_c128.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_spmv_c64.argtypes = [_i64, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.or_spmv_t_c128.argtypes = [_i64, _i64, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_l2g_update_f64.argtypes = [_i64, _i64, _i64, _vp, _vp, _vp, _vp]
        L.or_l2g_reverse_update_f64.argtypes = [_i64, _i64, _i64, _vp, _vp, _vp, _vp]
        L.or_cg_f64.argtypes = [_i64, _i64, _i32p, _i32p, _f64p, _f64p, ctypes.c_int,
                                ctypes.c_double, _vp, _f64p, _vp]
        L.or_cg_f64.restype = ctypes.c_int
        L.or_cg_f64_mt.argtypes = [_i64, _i64, _i32p, _i32p, _f64p, _f64p, ctypes.c_int,
                                   ctypes.c_double, _vp, _f64p, _vp, ctypes.c_int]
        L.or_cg_f64_mt.restype = ctypes.c_int
        L.or_cg_f32.argtypes = [_i64, _i32p, _i32p, _f32p, _f32p, ctypes.c_int,
                                ctypes.c_float, _f32p, _vp]
        L.or_cg_f32.restype = ctypes.c_int
        L.or_cg_f32_restart.argtypes = [_i64, _i32p, _i32p, _f32p, _f32p, ctypes.c_int,
                                        ctypes.c_float, _f32p, _vp, ctypes.c_int]
        L.or_cg_f32_restart.restype = ctypes.c_int
        L.or_gen_demo1d.argtypes = [_i64, ctypes.c_double, _vp, _vp, _vp]
        L.or_gen_demo1d.restype = _i64
        L.or_gen_stencil.argtypes = [ctypes.c_int, _i64, _i64, _i64, _i64,
                                     ctypes.c_double, ctypes.c_double, _vp, _vp, _vp]
        L.or_gen_stencil.restype = _i64
        L.or_diffusion_num_edges.argtypes = [_i64]
        L.or_diffusion_num_edges.restype = _i64
        L.or_gen_diffusion.argtypes = [_i64, _vp, _vp, _vp, _vp]
        L.or_gen_diffusion.restype = _i64
        L.or_diffusion_grad_theta.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p]
        L.or_adjoint_diffusion.argtypes = [_i64, _f64p, _f64p, _f64p, ctypes.c_int,
                                           ctypes.c_double, _f64p, _f64p, _f64p, _f64p]
        L.or_adjoint_diffusion.restype = ctypes.c_int
        L.or_uniform.argtypes = [ctypes.c_uint64, _i64, ctypes.c_double, ctypes.c_double, _f64p]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- RNG
GAMMA = np.uint64(0x9E3779B97F4A7C15)


def rand_u64(seed, n, start=0):
    """splitmix64 stream: value i = mix(seed + (i+1)*gamma), uint64 wraparound."""
    with np.errstate(over="ignore"):
        i = np.arange(start, start + n, dtype=np.uint64)
        z = np.uint64(seed) + (i + np.uint64(1)) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, n, lo=0.0, hi=1.0):
    u = (rand_u64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return lo + (hi - lo) * u


# ---------------------------------------------------------------- generators
class CSR:
    def __init__(self, nrows, ncols, rowptr, colidx, val):
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.rowptr, self.colidx, self.val = rowptr, colidx, val

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.colidx, self.rowptr), shape=(self.nrows, self.ncols))


def gen_demo1d(N, gamma=0.1):
    L = lib()
    nnz = L.or_gen_demo1d(N, gamma, None, None, None)
    rp = np.zeros(N + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz, np.float64)
    L.or_gen_demo1d(N, gamma, _ptr(rp), _ptr(ci), _ptr(v))
    return CSR(N, N, rp, ci, v)


def gen_stencil(kind, nx, ny, nz=1, ghost_planes=0, diag=None, off=-1.0):
    if diag is None:
        diag = {5: 4.0, 7: 6.0, 27: 26.0}[kind]
    L = lib()
    nnz = L.or_gen_stencil(kind, nx, ny, nz, ghost_planes, diag, off, None, None, None)
    M = nx * ny * nz
    rp = np.zeros(M + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz, np.float64)
    L.or_gen_stencil(kind, nx, ny, nz, ghost_planes, diag, off, _ptr(rp), _ptr(ci), _ptr(v))
    return CSR(M, M + nx * ny * ghost_planes, rp, ci, v)


def diffusion_num_edges(n):
    return int(lib().or_diffusion_num_edges(n))


def gen_diffusion(n, theta):
    L = lib()
    theta = np.ascontiguousarray(theta, np.float64)
    nnz = L.or_gen_diffusion(n, None, None, None, None)
    rp = np.zeros(n * n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz, np.float64)
    L.or_gen_diffusion(n, _ptr(theta), _ptr(rp), _ptr(ci), _ptr(v))
    return CSR(n * n, n * n, rp, ci, v)


# ---------------------------------------------------------------- kernels
def spmv(A, x, threads=1):
    y = np.empty(A.nrows, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    if threads == 1:
        lib().or_spmv_f64(A.nrows, A.rowptr, A.colidx, A.val, x, y)
    else:
        lib().or_spmv_f64_omp(A.nrows, A.rowptr, A.colidx, A.val, x, y, threads)
    return y


def spmv_t(A, x):
    y = np.empty(A.ncols, np.float64)
    lib().or_spmv_t_f64(A.nrows, A.ncols, A.rowptr, A.colidx, A.val,
                        np.ascontiguousarray(x, np.float64), y)
    return y


def spmv_f32(A, val32, x):
    y = np.empty(A.nrows, np.float32)
    lib().or_spmv_f32(A.nrows, A.rowptr, A.colidx, np.ascontiguousarray(val32, np.float32),
                      np.ascontiguousarray(x, np.float32), y)
    return y


def spmv_complex(A, values, x):
    """Complex product in the reference's row order; values and x complex128
    (or complex64 for single precision)."""
    if np.asarray(values).dtype == np.complex64:
        val = np.ascontiguousarray(values, np.complex64).view(np.float32)
        xx = np.ascontiguousarray(x, np.complex64).view(np.float32)
        y = np.empty(2 * A.nrows, np.float32)
        lib().or_spmv_c64(A.nrows, A.rowptr, A.colidx, val, xx, y)
        return y.view(np.complex64)
    val = np.ascontiguousarray(values, np.complex128).view(np.float64)
    xx = np.ascontiguousarray(x, np.complex128).view(np.float64)
    y = np.empty(2 * A.nrows, np.float64)
    lib().or_spmv_c128(A.nrows, A.rowptr, A.colidx, val, xx, y)
    return y.view(np.complex128)


def spmv_t_complex(A, values, x):
    val = np.ascontiguousarray(values, np.complex128).view(np.float64)
    xx = np.ascontiguousarray(x, np.complex128).view(np.float64)
    y = np.empty(2 * A.ncols, np.float64)
    lib().or_spmv_t_c128(A.nrows, A.ncols, A.rowptr, A.colidx, val, xx, y)
    return y.view(np.complex128)


def l2g_update(local_size, num_ghosts, indexbuf, v, halo):
    """In place on v (float64); returns the packed send buffer."""
    indexbuf = np.ascontiguousarray(indexbuf, np.int32)
    send = np.empty(len(indexbuf), np.float64)
    halo = None if halo is None else np.ascontiguousarray(halo, np.float64)
    lib().or_l2g_update_f64(local_size, num_ghosts, len(indexbuf), _ptr(indexbuf),
                            _ptr(v), _ptr(halo), _ptr(send))
    return send


def l2g_reverse_update(local_size, num_ghosts, indexbuf, v, incoming):
    """In place on v; returns the ghost values sent to the owners."""
    indexbuf = np.ascontiguousarray(indexbuf, np.int32)
    out = np.empty(num_ghosts, np.float64)
    incoming = np.ascontiguousarray(incoming, np.float64)
    lib().or_l2g_reverse_update_f64(local_size, num_ghosts, len(indexbuf), _ptr(indexbuf),
                                    _ptr(v), _ptr(incoming), _ptr(out))
    return out


def cg(A, b, kmax, rtol, halo=None, threads=None):
    """spmv::cg restated; returns (x[ncols], k, hist[0..k] of squared norms).
    threads=None: one thread, sequential dot products.  threads=n: the product,
    vector updates and dot products split over n OpenMP threads, the dot
    products summed in a fixed 64-chunk order (result independent of n; it
    differs from the sequential one only by the dot products' rounding)."""
    x = np.empty(A.ncols, np.float64)
    hist = np.zeros(kmax + 1, np.float64)
    halo = None if halo is None else np.ascontiguousarray(halo, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if threads is not None:
        k = lib().or_cg_f64_mt(A.nrows, A.ncols, A.rowptr, A.colidx, A.val, b, kmax, rtol,
                               _ptr(halo), x, _ptr(hist), int(threads))
    else:
        k = lib().or_cg_f64(A.nrows, A.ncols, A.rowptr, A.colidx, A.val, b, kmax, rtol,
                            _ptr(halo), x, _ptr(hist))
    return x, k, hist[: k + 1].copy()


def cg_f32(A, val32, b, kmax, rtol, restarts=0):
    """FP32 CG; restarts > 0 adds the true-residual check of the batched
    solver (b - A x in double when the recursive residual meets rtol, restart
    from x if the true one does not, at most `restarts` times)."""
    x = np.empty(A.nrows, np.float32)
    hist = np.zeros(kmax + 1, np.float32)
    k = lib().or_cg_f32_restart(A.nrows, A.rowptr, A.colidx,
                                np.ascontiguousarray(val32, np.float32),
                                np.ascontiguousarray(b, np.float32), kmax, rtol, x, _ptr(hist),
                                int(restarts))
    return x, k, hist[: k + 1].copy()


def adjoint_diffusion(n, theta, b, x_target, kmax, rtol):
    N = n * n
    x = np.empty(N)
    lam = np.empty(N)
    g = np.empty(diffusion_num_edges(n))
    loss = np.zeros(1)
    its = lib().or_adjoint_diffusion(n, np.ascontiguousarray(theta, np.float64),
                                     np.ascontiguousarray(b, np.float64),
                                     np.ascontiguousarray(x_target, np.float64),
                                     kmax, rtol, x, lam, g, loss)
    return dict(x=x, lam=lam, grad_theta=g, grad_b=lam.copy(), loss=float(loss[0]), its=its)


def diffusion_grad_theta(n, A, G):
    g = np.empty(diffusion_num_edges(n))
    lib().or_diffusion_grad_theta(n, A.rowptr, A.colidx, np.ascontiguousarray(G, np.float64), g)
    return g
