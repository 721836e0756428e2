"""Python binding of libchuckle_hip.so, the MI355X (gfx950) implementation of
Chuckle's file-to-file search hot path.

The compute runs in hand-written HIP kernels behind the C ABI declared in
``include/chuckle_hip.h``; this module only marshals numpy arrays through
ctypes.  There is no CPU fallback: if the library is missing, importing the
handle types raises ``RuntimeError`` telling the caller to build it
(``python -m riggle_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CHUCKLE_HIP_LIB") or os.path.join(HERE, "libchuckle_hip.so")

CHK_OK = 0
MAX_NOALLOC = 1 << 20  # CHK_NOALLOC_MAX_INTERVALS
_ERRORS = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "I/O error",
           5: "not found", 6: "internal error"}

_lib = None


class ChuckleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load libchuckle_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m riggle_amd.build`"
                               " (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    L.chk_version.restype = C.c_char_p
    L.chk_last_error.restype = C.c_char_p
    L.chk_device_count.argtypes = [C.POINTER(C.c_int)]
    L.chk_index_create.argtypes = [i32, u32, u32, u64, vp, vp, vp, vp, C.POINTER(vp)]
    L.chk_index_free.argtypes = [vp]
    L.chk_index_info.argtypes = [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u64), C.POINTER(u64)]
    L.chk_query_create.argtypes = [vp, u32, u64, vp, vp, vp, vp, C.POINTER(vp)]
    L.chk_query_free.argtypes = [vp]
    L.chk_query_reset.argtypes = [vp, u32, u64, vp, vp, vp, vp]
    L.chk_moments_widths.argtypes = [vp, u32, vp, vp, vp, C.POINTER(u32), C.POINTER(u64), vp,
                                     C.POINTER(u64)]
    L.chk_query_info.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(u64), C.POINTER(C.c_uint32),
                                 C.POINTER(u64)]
    L.chk_sweep.argtypes = [vp, vp, vp, vp]
    L.chk_sweep_async.argtypes = [vp, vp, vp, vp]
    L.chk_dev_alloc.argtypes = [vp, u64, C.POINTER(vp)]
    L.chk_dev_free.argtypes = [vp, vp]
    L.chk_dev_to_host.argtypes = [vp, vp, vp, u64]
    L.chk_sync.argtypes = [vp]
    L.chk_profile_enable.argtypes = [vp, i32]
    L.chk_profile_only.argtypes = [vp, C.c_char_p]
    L.chk_profile_read.argtypes = [vp, C.c_char_p, C.POINTER(u64), C.POINTER(C.c_double)]
    L.chk_calibrate_copy.argtypes = [vp, u64]
    L.chk_calibrate_read.argtypes = [vp, u32, C.POINTER(C.c_double), C.POINTER(u64)]
    L.chk_tune.argtypes = [vp, C.c_char_p, C.c_int64]
    L.chk_index_set_genome.argtypes = [vp, vp]
    L.chk_moments_upload.argtypes = [vp, vp, vp, u64, vp, vp]
    L.chk_pvalues_async.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.chk_moments_build.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.chk_moments_download.argtypes = [vp, C.POINTER(u64), vp, vp, vp, vp]
    L.chk_query_stats.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.chk_tune_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
    L.chk_autotune.argtypes = [vp, vp, vp, vp, i32, C.POINTER(C.c_double), C.c_char_p, u64]
    L.chk_graph_create.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]
    L.chk_graph_launch.argtypes = [vp]
    L.chk_graph_free.argtypes = [vp]
    dp_ = C.POINTER(C.c_double)
    L.chk_server_create.argtypes = [vp, u32, u32, u64, i32, C.POINTER(vp)]
    L.chk_server_submit.argtypes = [vp, u64, vp, vp, vp, C.POINTER(u64)]
    L.chk_server_wait.argtypes = [vp, u64, vp, vp, vp, vp, vp, dp_]
    L.chk_server_stats.argtypes = [vp, C.POINTER(u64), dp_, dp_, dp_, dp_, dp_]
    L.chk_server_reset_stats.argtypes = [vp]
    L.chk_server_reserve.argtypes = [vp, u64]
    L.chk_server_free.argtypes = [vp]
    L.chk_sweep_noalloc.argtypes = [vp, u64, vp, vp, vp, vp, vp]
    cpp = C.POINTER(C.c_char_p)
    L.chk_index_open.argtypes = [C.c_char_p, i32, C.POINTER(vp)]
    L.chk_index_shard_name.argtypes = [vp, u32, C.c_char_p, u64, vp]
    L.chk_index_file_name.argtypes = [vp, u32, C.c_char_p, u64, vp]
    L.chk_index_set_names.argtypes = [vp, cpp, cpp]
    L.chk_query_open.argtypes = [vp, u32, cpp, vp, C.POINTER(vp)]
    L.chk_db_add.argtypes = [C.c_char_p, C.c_char_p, u64, i32, cpp, u32, i32, i32]
    L.chk_write_results.argtypes = [C.c_char_p, u32, cpp, u32, cpp, vp, vp, vp, vp, vp, i32]


def _check(code):
    if code != CHK_OK:
        raise ChuckleError(code, lib().chk_last_error().decode())


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and len(a) else None


def version():
    return lib().chk_version().decode()


def device_count():
    n = C.c_int()
    _check(lib().chk_device_count(C.byref(n)))
    return n.value


def _cstrs(names):
    arr = (C.c_char_p * max(1, len(names)))()
    for i, n in enumerate(names):
        arr[i] = n.encode()
    return arr


def db_add(input_path, db_path, batch_size=0, stats=False, chroms=(), sorted_sids=False, device=0):
    """`chuckle add`: BED files of input_path into the database at db_path."""
    _check(lib().chk_db_add(str(input_path).encode(), str(db_path).encode(), int(batch_size),
                            1 if stats else 0, _cstrs(list(chroms)), len(chroms),
                            1 if sorted_sids else 0, device))


def write_results(path, query_names, db_names, counts, bp=None, p=None, llr=None, flags=None,
                  stats=False):
    """StatsOutput JSON of a result matrix [len(query_names) x len(db_names)]."""
    arrs = [np.ascontiguousarray(a) if a is not None else None for a in (counts, bp, p, llr, flags)]
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    _check(lib().chk_write_results(str(path).encode(), len(query_names), _cstrs(query_names),
                                   len(db_names), _cstrs(db_names), *[ptr(a) for a in arrs],
                                   1 if stats else 0))


class Index:
    """HBM-resident DB index (per-shard start/end streams + file ids)."""

    @classmethod
    def open(cls, db_path, device=0):
        """Index of a database directory in the reference's format."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        _check(lib().chk_index_open(str(db_path).encode(), device, C.byref(h)))
        self._h = h
        i = self.info()
        self.nshards, self.nfiles, self.n = i["nshards"], i["nfiles"], i["n"]
        return self

    def names(self):
        """(shard names, file names) of an opened or named index."""
        buf = C.create_string_buffer(4096)
        out = []
        for fn, k in ((lib().chk_index_shard_name, self.nshards), (lib().chk_index_file_name, self.nfiles)):
            cur = []
            for i in range(k):
                _check(fn(self._h, i, buf, len(buf), None))
                cur.append(buf.value.decode())
            out.append(cur)
        return tuple(out)

    def set_names(self, shard_names, file_names):
        assert len(shard_names) == self.nshards and len(file_names) == self.nfiles
        _check(lib().chk_index_set_names(self._h, _cstrs(shard_names), _cstrs(file_names)))

    def query_files(self, paths):
        """Parse BED / BED.gz files (host threads) into one query batch;
        returns (Query or None when every file is empty, row per path or -1)."""
        rows = np.zeros(max(1, len(paths)), np.uint32)
        h = C.c_void_p()
        _check(lib().chk_query_open(self._h, len(paths), _cstrs([str(p) for p in paths]),
                                    rows.ctypes.data_as(C.c_void_p), C.byref(h)))
        rows = rows[:len(paths)].astype(np.int64)
        rows[rows == 0xFFFFFFFF] = -1
        if not h.value:
            return None, rows
        q = Query.__new__(Query)
        q._h, q._index, q.nfiles = h, self, int((rows >= 0).sum())
        q.n = None
        return q, rows

    def __init__(self, shard, start, end, fid, nshards, nfiles, device=0):
        shard, start, end, fid = _u32(shard), _u32(start), _u32(end), _u32(fid)
        self._keep = (shard, start, end, fid)
        h = C.c_void_p()
        _check(lib().chk_index_create(device, nshards, nfiles, len(start), _ptr(shard),
                                      _ptr(start), _ptr(end), _ptr(fid), C.byref(h)))
        self._h = h
        self._keep = None
        self.nshards, self.nfiles, self.n = int(nshards), int(nfiles), len(start)

    @property
    def handle(self):
        return self._h

    def info(self):
        a, b, n, by = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(lib().chk_index_info(self._h, C.byref(a), C.byref(b), C.byref(n), C.byref(by)))
        return {"nshards": a.value, "nfiles": b.value, "n": n.value, "device_bytes": by.value}

    def query(self, shard, start, end, qfile=None, nfiles=1):
        return Query(self, shard, start, end, qfile, nfiles)

    def sweep(self, query):
        """Exact (counts, bp) uint64 matrices [query files x DB files]."""
        cnt = np.zeros((query.nfiles, self.nfiles), np.uint64)
        bp = np.zeros((query.nfiles, self.nfiles), np.uint64)
        _check(lib().chk_sweep(self._h, query.handle, cnt.ctypes.data_as(C.c_void_p),
                               bp.ctypes.data_as(C.c_void_p)))
        return cnt, bp

    def sweep_noalloc(self, shard, start, end, d_counts, d_bp):
        """One query file swept without any device allocation (static scratch;
        at most MAX_NOALLOC intervals) into device buffers of nfiles u64."""
        shard, start, end = _u32(shard), _u32(start), _u32(end)
        _check(lib().chk_sweep_noalloc(self._h, len(start), _ptr(shard), _ptr(start), _ptr(end),
                                       d_counts.ptr, d_bp.ptr))

    def sweep_async(self, query, d_counts, d_bp):
        _check(lib().chk_sweep_async(self._h, query.handle, d_counts.ptr, d_bp.ptr))

    def sync(self):
        _check(lib().chk_sync(self._h))

    def profile(self, on=True, only=()):
        """Record kernel timings (only the named regions when `only` is given)."""
        _check(lib().chk_profile_only(self._h, ",".join(only).encode()))
        _check(lib().chk_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, name):
        n, ms = C.c_uint64(), C.c_double()
        _check(lib().chk_profile_read(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def set_genome(self, shard_size):
        """Per-shard length in bp (0 = not in the genome table)."""
        a = np.ascontiguousarray(shard_size, np.uint64)
        assert len(a) == self.nshards
        _check(lib().chk_index_set_genome(self._h, a.ctypes.data_as(C.c_void_p)))

    def upload_moments(self, slot_off, slot_len, mean, var):
        so = np.ascontiguousarray(slot_off, np.int64)
        sl = np.ascontiguousarray(slot_len, np.uint32)
        m = np.ascontiguousarray(mean, np.float64)
        v = np.ascontiguousarray(var, np.float64)
        assert len(so) == self.nfiles * self.nshards == len(sl)
        _check(lib().chk_moments_upload(self._h, so.ctypes.data_as(C.c_void_p),
                                        sl.ctypes.data_as(C.c_void_p), len(m),
                                        m.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p)))

    def build_moments(self):
        """Compute the compact moment tables on the device; returns
        (signals, signals with exact integer autocorrelation)."""
        rows, ex = C.c_int32(), C.c_int32()
        _check(lib().chk_moments_build(self._h, C.byref(rows), C.byref(ex)))
        return rows.value, ex.value

    def moments_widths(self, widths, lags=False):
        """FP32 lag-limited route: (mean, var) [files x shards x widths] at the
        given block sizes (<= 1024 bins), plus {"inexact_rows", "bins"} and,
        with lags=True, the autocorrelation lags [files x shards x 1024]."""
        w = np.ascontiguousarray(widths, np.uint32)
        shp = (self.nfiles, self.nshards, len(w))
        m = np.zeros(shp, np.float64)
        v = np.zeros(shp, np.float64)
        lg = np.zeros((self.nfiles, self.nshards, 1024), np.float64) if lags else None
        bad, bins, ex = C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(lib().chk_moments_widths(self._h, len(w), w.ctypes.data_as(C.c_void_p),
                                        m.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                                        C.byref(bad), C.byref(bins),
                                        lg.ctypes.data_as(C.c_void_p) if lags else None, C.byref(ex)))
        info = {"inexact_rows": bad.value, "bins": bins.value, "exact_widths": ex.value}
        return (m, v, info, lg) if lags else (m, v, info)

    def download_moments(self):
        """(slot_off [F*S], slot_len [F*S], mean, var) of the installed tables."""
        tot = C.c_uint64()
        _check(lib().chk_moments_download(self._h, C.byref(tot), None, None, None, None))
        ns = self.nfiles * self.nshards
        so = np.zeros(ns, np.int64)
        sl = np.zeros(ns, np.uint32)
        m = np.zeros(max(tot.value, 1), np.float64)
        v = np.zeros(max(tot.value, 1), np.float64)
        _check(lib().chk_moments_download(self._h, C.byref(tot), so.ctypes.data_as(C.c_void_p),
                                          sl.ctypes.data_as(C.c_void_p), m.ctypes.data_as(C.c_void_p),
                                          v.ctypes.data_as(C.c_void_p)))
        return so, sl, m[:tot.value], v[:tot.value]

    def stats(self, query):
        """Sweep + p-values: dict of [query files x DB files] arrays."""
        sh = (query.nfiles, self.nfiles)
        out = {"counts": np.zeros(sh, np.uint64), "bp": np.zeros(sh, np.uint64),
               "p": np.zeros(sh, np.float64), "llr": np.zeros(sh, np.float64),
               "flags": np.zeros(sh, np.uint8)}
        _check(lib().chk_query_stats(self._h, query.handle,
                                     *[out[k].ctypes.data_as(C.c_void_p)
                                       for k in ("counts", "bp", "p", "llr", "flags")]))
        return out

    def pvalues_async(self, query, d_counts, d_bp, d_p, d_llr, d_flags):
        _check(lib().chk_pvalues_async(self._h, query.handle, d_counts.ptr, d_bp.ptr, d_p.ptr,
                                       d_llr.ptr, d_flags.ptr))

    def tune(self, key, value):
        _check(lib().chk_tune(self._h, key.encode(), int(value)))

    def tune_get(self, key):
        v = C.c_int64()
        _check(lib().chk_tune_get(self._h, key.encode(), C.byref(v)))
        return v.value

    def autotune(self, query, d_counts, d_bp, reps=3):
        """Pick the sweep knobs for this batch by timing; returns
        (best ms per sweep, chosen knob values, report lines)."""
        best = C.c_double()
        buf = C.create_string_buffer(16384)
        _check(lib().chk_autotune(self._h, query.handle, d_counts.ptr, d_bp.ptr, int(reps),
                                  C.byref(best), buf, len(buf)))
        keys = ("sweep.sparse_max", "sweep.items", "sweep.chunk", "sweep.blocks_per_cu")
        return best.value, {k: self.tune_get(k) for k in keys}, buf.value.decode().splitlines()

    def graph(self, query, d_counts, d_bp, d_p=None, d_llr=None, d_flags=None):
        """Capture the batch's sweep (and p-values) as one replayable HIP graph."""
        return Graph(self, query, d_counts, d_bp, d_p, d_llr, d_flags)

    def calibrate_read(self, reps=3):
        """(best ms, bytes) of a read-only pass over the dense streams."""
        ms, by = C.c_double(), C.c_uint64()
        _check(lib().chk_calibrate_read(self._h, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def calibrate_copy(self, nbytes):
        _check(lib().chk_calibrate_copy(self._h, nbytes))

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def close(self):
        if getattr(self, "_h", None):
            lib().chk_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Query:
    """A batch of query files uploaded to the device."""

    def __init__(self, index, shard, start, end, qfile=None, nfiles=1):
        shard, start, end = _u32(shard), _u32(start), _u32(end)
        qf = _u32(qfile) if qfile is not None else None
        h = C.c_void_p()
        _check(lib().chk_query_create(index.handle, nfiles, len(start), _ptr(shard), _ptr(start),
                                      _ptr(end), _ptr(qf), C.byref(h)))
        self._h = h
        self._index = index
        self.nfiles = int(nfiles)
        self.n = len(start)

    @property
    def handle(self):
        return self._h

    def reset(self, shard, start, end, qfile=None, nfiles=1):
        """Replace the batch, reusing this handle's buffers."""
        shard, start, end = _u32(shard), _u32(start), _u32(end)
        qf = _u32(qfile) if qfile is not None else None
        _check(lib().chk_query_reset(self._h, nfiles, len(start), _ptr(shard), _ptr(start), _ptr(end),
                                     _ptr(qf)))
        self.nfiles, self.n = int(nfiles), len(start)

    def info(self):
        """Sweep plan under the index's current tuning: files and intervals
        kept, and the files (and intervals) on the dense merge-path sweep."""
        nf, n, df = C.c_uint32(), C.c_uint64(), C.c_uint32()
        dn = C.c_uint64()
        _check(lib().chk_query_info(self._h, C.byref(nf), C.byref(n), C.byref(df), C.byref(dn)))
        return {"nfiles": nf.value, "n": n.value, "dense_files": df.value,
                "dense_intervals": dn.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().chk_query_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over an index: submit() query files from any thread,
    wait() for each result row; stats() gives latency percentiles."""

    def __init__(self, index, slots=2, max_batch_files=64, hbm_budget=0, pvalues=False):
        h = C.c_void_p()
        _check(lib().chk_server_create(index.handle, slots, max_batch_files, hbm_budget,
                                       1 if pvalues else 0, C.byref(h)))
        self._h, self._index, self.pvalues = h, index, pvalues

    def submit(self, shard, start, end):
        shard, start, end = _u32(shard), _u32(start), _u32(end)
        t = C.c_uint64()
        _check(lib().chk_server_submit(self._h, len(start), _ptr(shard), _ptr(start), _ptr(end),
                                       C.byref(t)))
        return t.value

    def wait(self, ticket):
        F = self._index.nfiles
        out = {"counts": np.zeros(F, np.uint64), "bp": np.zeros(F, np.uint64),
               "p": np.zeros(F, np.float64), "llr": np.zeros(F, np.float64),
               "flags": np.zeros(F, np.uint8)}
        lat = C.c_double()
        _check(lib().chk_server_wait(self._h, ticket,
                                     *[out[k].ctypes.data_as(C.c_void_p)
                                       for k in ("counts", "bp", "p", "llr", "flags")],
                                     C.byref(lat)))
        out["latency_ms"] = lat.value
        return out

    def stats(self):
        n = C.c_uint64()
        v = [C.c_double() for _ in range(5)]
        _check(lib().chk_server_stats(self._h, C.byref(n), *[C.byref(x) for x in v]))
        keys = ("p50_ms", "p99_ms", "mean_ms", "swept_per_s", "files_per_batch")
        return {"completed": n.value, **{k: x.value for k, x in zip(keys, v)}}

    def reset_stats(self):
        _check(lib().chk_server_reset_stats(self._h))

    def reserve(self, max_batch_intervals):
        """Grow every slot for batches of up to this many intervals (idle
        server only), so that serving allocates nothing afterwards."""
        _check(lib().chk_server_reserve(self._h, int(max_batch_intervals)))

    def close(self):
        if getattr(self, "_h", None):
            lib().chk_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """One captured sweep (+ p-values) of a query batch; launch() replays it on
    the index's stream."""

    def __init__(self, index, query, d_counts, d_bp, d_p=None, d_llr=None, d_flags=None):
        h = C.c_void_p()
        ptr = lambda b: b.ptr if b is not None else None
        _check(lib().chk_graph_create(index.handle, query.handle, d_counts.ptr, d_bp.ptr,
                                      ptr(d_p), ptr(d_llr), ptr(d_flags), C.byref(h)))
        self._h, self._index, self._query = h, index, query
        # the captured kernels write these buffers on every launch: keep them
        # alive for as long as the graph is
        self._bufs = (d_counts, d_bp, d_p, d_llr, d_flags)

    def launch(self):
        _check(lib().chk_graph_launch(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().chk_graph_free(self._h)
            self._h = None
        self._bufs = ()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    def __init__(self, index, nbytes):
        p = C.c_void_p()
        _check(lib().chk_dev_alloc(index.handle, nbytes, C.byref(p)))
        self.ptr, self.nbytes, self._index = p, int(nbytes), index

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().chk_dev_to_host(self._index.handle, out.ctypes.data_as(C.c_void_p), self.ptr,
                                     out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().chk_dev_free(self._index.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
