"""Autotuner of the launch parameters.

* ``tune_batch``  -- workgroup size and lanes per row (rows per wavefront =
  64 / lanes) of the persistent batched CG kernel, timed on the batch itself;
* ``tune_spmv``   -- SpMV kernel variant (staging chunk, vector width, loads
  in flight, pipelining, or the sliced row-pattern layout) of a matrix.

Results are cached in a JSON file keyed by a signature of the problem
(``$SPMV_AMD_TUNE_CACHE``, default ``~/.cache/spmv_amd/tune.json``), so a
second run with the same shapes applies the stored choice without timing.
"""
import hashlib
import json
import os

import numpy as np

BATCH_BLOCKS = (256, 512, 1024)
BATCH_LANES = (0, 2, 4, 8)
BATCH_ROWS_PER_MEMBER = (4096, 16384, 65536)


def cache_path():
    return os.environ.get("SPMV_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "spmv_amd", "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    data = _load()
    data[key] = value
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _device_name():
    import torch
    return torch.cuda.get_device_name(0)


def _time_ms(fn, reps):
    import torch
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def batch_signature(batch, max_its, rtol):
    h = hashlib.sha1(np.asarray(batch.sys_row, np.int64).tobytes()).hexdigest()[:16]
    return f"batch:{_device_name()}:{batch.nsys}:{batch.nrows}:{batch.nnz}:{batch.nrhs}:{h}:{max_its}:{rtol:g}"


def tune_batch(batch, b, max_its=10000, rtol=1e-6, reps=2, blocks=BATCH_BLOCKS,
               lanes=BATCH_LANES, rows_per_member=BATCH_ROWS_PER_MEMBER, use_cache=True):
    """Times every (workgroup size, lanes per row, rows per team member)
    triple on the batch solve, applies the fastest and returns (best, table)."""
    import torch
    key = batch_signature(batch, max_its, rtol)
    if use_cache:
        hit = _load().get(key)
        if hit:
            batch.set_config(hit["block"], hit["lanes_per_row"], hit["rows_per_member"])
            return hit, None
    x = torch.empty_like(b)
    table = []
    for bs in blocks:
        for g in lanes:
            for rpm in rows_per_member:
                batch.set_config(bs, g, rpm)
                batch.solve(b, x, max_its, rtol, iterations=False)  # capture + warm up
                ms = _time_ms(lambda: batch.solve(b, x, max_its, rtol, iterations=False), reps)
                table.append({"block": bs, "lanes_per_row": g, "rows_per_member": rpm, "ms": ms})
    best = min(table, key=lambda r: r["ms"])
    batch.set_config(best["block"], best["lanes_per_row"], best["rows_per_member"])
    _store(key, best)
    return best, table


def spmv_signature(A):
    from . import lib
    return (f"spmv:{_device_name()}:{lib().spmvx_num_kernels()}:{A.dtype}:"
            f"{A.nrows}:{A.ncols}:{A.nnz}")


def tune_spmv(A, x, reps=10, use_cache=True):
    """Times every SpMV kernel variant of matrix A (all give identical
    results), applies the fastest and returns (best, table)."""
    import torch
    from . import lib
    from . import SpmvError
    key = spmv_signature(A)
    if use_cache:
        hit = _load().get(key)
        if hit:
            # the signature is the shape only: a choice stored for another
            # matrix of this shape (a layout this one does not fit) or by a
            # build whose variant table differed is tuned again
            try:
                A.set_kernel(hit["variant"])
                if hit.get("kernel") == A.kernel()[1]:
                    return hit, None
            except SpmvError:
                pass
    y = torch.empty(A.nrows, dtype=A.vector_dtype(), device=x.device)
    table = []
    for v in range(lib().spmvx_num_kernels()):
        try:
            A.set_kernel(v)
        except SpmvError:  # a layout the matrix does not fit (sliced row patterns)
            continue
        A.mult(x, out=y)
        ms = _time_ms(lambda: A.mult(x, out=y), reps)
        table.append({"variant": v, "ms": ms})
    best = dict(min(table, key=lambda r: r["ms"]))
    A.set_kernel(best["variant"])
    best["kernel"] = A.kernel()[1]
    _store(key, best)
    return best, table
