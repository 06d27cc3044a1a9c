"""Python host mirror of the graph-bit-ops engine (ctypes over libgbo.so).

The reference (carolc0708/graph-bit-ops) exposes its path at source level:
``csr2bsr_batch_<N>(h_csrRowPtr, h_csrColInd, nrows, ncols, nnz, d_bsrRowPtr,
d_bsrColInd, d_bsrVal, blocksize, nblockrows, nblocks)``
(algorithm/csr2bsr_batch_bsrbmv.cu:180,406,621,786) followed by bmv/bmm kernel
launches, all driven from C++ ``main`` functions.  This module binds the C ABI
that replaces those entry points (include/gbo.h) and keeps the reference's
names and argument meaning where one exists (``csr2bsr_batch``,
``bsrbmv``/``bsrbmm`` operations, per-algorithm drivers).

Only numpy and ctypes are needed; device memory is managed through the
library's own HIP wrappers (``DeviceBuffer``).  Every call goes to the HIP
kernels in libgbo.so; there is no CPU fallback -- a missing library raises.
"""
from __future__ import annotations

import ctypes
import os
import re
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgbo.so")
_LIB = None

GBO_OK = 0
ERROR_NAMES = {
    -1: "GBO_EINVAL", -2: "GBO_ENOMEM", -3: "GBO_EDEVICE", -4: "GBO_EIO",
    -5: "GBO_EFORMAT", -6: "GBO_ERANGE", -7: "GBO_ENOCONV", -8: "GBO_EUNSUPPORTED",
}
BLOCK_SIZES = (8, 16, 32, 64)
WORD_DTYPE = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}
INF_U32 = 0xFFFFFFFF


class Tuning(ctypes.Structure):
    """Mirror of ``gbo_tuning``."""

    _fields_ = [("bs", c_int32), ("chunk", c_int32), ("variant", c_int32), ("ms", c_float)]


class GboError(RuntimeError):
    """A negative GBO_E* status returned through the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


class BsrStruct(ctypes.Structure):
    """Mirror of ``gbo_bsr`` in include/gbo.h."""

    _fields_ = [
        ("bs", c_int32), ("dir", c_int32), ("nrows", c_int32), ("ncols", c_int32),
        ("nblockrows", c_int32), ("nblockcols", c_int32), ("nblocks", c_int64),
        ("rowptr", c_void_p), ("colind", c_void_p), ("tiles", c_void_p),
        ("plan", c_void_p), ("plan_chunk", c_int32), ("plan_nchunks", c_int32),
        ("plan_variant", c_int32),
    ]


def header_symbols(header_path: str | None = None) -> list[str]:
    """Names of every function declared in include/gbo.h."""
    if header_path is None:
        header_path = os.path.join(_HERE, "..", "include", "gbo.h")
    text = open(header_path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gbo_[a-z0-9_]+)\s*\(", text)))


def _declare(L):
    P = c_void_p
    i32p = POINTER(c_int32)
    i64p = POINTER(c_int64)
    sigs = {
        "gbo_last_error": ([], c_char_p),
        "gbo_version": ([], c_char_p),
        "gbo_device_count": ([POINTER(c_int)], c_int),
        "gbo_set_device": ([c_int], c_int),
        "gbo_malloc": ([POINTER(c_void_p), c_size_t], c_int),
        "gbo_free": ([P], c_int),
        "gbo_memcpy_h2d": ([P, P, c_size_t], c_int),
        "gbo_memcpy_d2h": ([P, P, c_size_t], c_int),
        "gbo_memcpy_d2d": ([P, P, c_size_t], c_int),
        "gbo_memset": ([P, c_int, c_size_t], c_int),
        "gbo_synchronize": ([], c_int),
        "gbo_stream_create": ([POINTER(c_void_p)], c_int),
        "gbo_stream_destroy": ([P], c_int),
        "gbo_stream_synchronize": ([P], c_int),
        "gbo_event_create": ([POINTER(c_void_p)], c_int),
        "gbo_event_destroy": ([P], c_int),
        "gbo_event_record": ([P, P], c_int),
        "gbo_event_elapsed": ([POINTER(c_float), P, P], c_int),
        "gbo_free_host": ([P], None),
        "gbo_stream_copy": ([P, P, c_size_t, P], c_int),
        "gbo_stream_read": ([P, c_size_t, c_uint32, P, P], c_int),
        "gbo_gather_probe": ([P, c_int, P, c_int64, c_uint32, P, P], c_int),
        "gbo_mtx_read": ([c_char_p, c_int, c_int, i32p, i32p, i64p, POINTER(P), POINTER(P), POINTER(P)], c_int),
        "gbo_mtx_read_cached": ([c_char_p, c_int, c_int, POINTER(c_int), i32p, i32p, i64p, POINTER(P), POINTER(P),
                                 POINTER(P)], c_int),
        "gbo_mtx_cache_path": ([c_char_p, c_int, c_int, c_char_p, c_size_t], c_int),
        "gbo_mtx_write": ([c_char_p, c_int32, c_int32, P, P], c_int),
        "gbo_gen_er": ([c_int32, c_double, c_uint64, i64p, POINTER(P), POINTER(P)], c_int),
        "gbo_gen_er_m": ([c_int32, c_int64, c_uint64, i64p, POINTER(P), POINTER(P)], c_int),
        "gbo_gen_rmat": ([c_int32, c_int32, c_double, c_double, c_double, c_uint64, i64p, POINTER(P), POINTER(P)], c_int),
        "gbo_gen_king": ([c_int32, i64p, POINTER(P), POINTER(P)], c_int),
        "gbo_gen_edge_weights": ([c_int32, P, P, c_uint32, c_uint32, c_uint64, P], c_int),
        "gbo_csr_transpose": ([c_int32, c_int32, P, P, POINTER(P), POINTER(P)], c_int),
        "gbo_csr_tril": ([c_int32, P, P, c_int, i64p, POINTER(P), POINTER(P)], c_int),
        "gbo_csr2bsr_nnzb": ([c_int, P, P, c_int32, c_int32, i64p], c_int),
        "gbo_csr2bsr_batch": ([c_int, c_int, P, P, c_int32, c_int32, c_int64, P, P, P, c_int32, c_int64, P], c_int),
        "gbo_csr2bsr_nnzb_dev": ([c_int, P, P, c_int32, c_int32, P, i64p, P], c_int),
        "gbo_csr2bsr_fill_dev": ([c_int, c_int, P, P, c_int32, c_int32, P, P, P, P], c_int),
        "gbo_bsr_create_dev": ([c_int, c_int, P, P, c_int32, c_int32, POINTER(BsrStruct), P], c_int),
        "gbo_bsr_create": ([c_int, c_int, P, P, c_int32, c_int32, POINTER(BsrStruct), P], c_int),
        "gbo_bsr_free": ([POINTER(BsrStruct)], None),
        "gbo_bsr_plan": ([POINTER(BsrStruct), c_int32, P], c_int),
        "gbo_bsr_plan_free": ([POINTER(BsrStruct)], None),
        "gbo_bsr_plan_ex": ([POINTER(BsrStruct), c_int32, c_int32, P], c_int),
        "gbo_csr2bsr_values_u32": ([POINTER(BsrStruct), P, P, P, P, P, P], c_int),
        "gbo_pack_vec": ([c_int, P, P, c_int32, P], c_int),
        "gbo_bsrbmv_count": ([POINTER(BsrStruct), P, P, c_int, P], c_int),
        "gbo_bsrbmv_bfs_step": ([POINTER(BsrStruct), P, P, P, P, P], c_int),
        "gbo_bsrbmv_minplus": ([POINTER(BsrStruct), P, P, P, P, P], c_int),
        "gbo_bsrbmv_sum_f32": ([POINTER(BsrStruct), P, P, P], c_int),
        "gbo_bsrbmv_min_i32": ([POINTER(BsrStruct), P, P, P], c_int),
        "gbo_bsrbmm_tc": ([POINTER(BsrStruct), POINTER(BsrStruct), P, P], c_int),
        "gbo_bsrbmm_sum": ([POINTER(BsrStruct), POINTER(BsrStruct), P, P], c_int),
        "gbo_bsrbmm_symbolic": ([POINTER(BsrStruct), POINTER(BsrStruct), P, i64p, P], c_int),
        "gbo_bsrbmm_numeric": ([POINTER(BsrStruct), POINTER(BsrStruct), P, c_int64, P, P, P, i64p, P], c_int),
        "gbo_bsrbmm_values": ([POINTER(BsrStruct), POINTER(BsrStruct), POINTER(BsrStruct), P, P, P], c_int),
        "gbo_bsrbmm": ([POINTER(BsrStruct), POINTER(BsrStruct), POINTER(BsrStruct), P], c_int),
        "gbo_bsr_bit_offsets": ([POINTER(BsrStruct), P, i64p, P], c_int),
        "gbo_bfs": ([POINTER(BsrStruct), c_int32, P, P, c_int32, i32p, P], c_int),
        "gbo_bfs_noalloc": ([POINTER(BsrStruct), c_int32, P, P, c_int32, i32p, P], c_int),
        "gbo_bfs_do": ([POINTER(BsrStruct), POINTER(BsrStruct), c_int32, P, P, c_int32, i32p, P], c_int),
        "gbo_sssp": ([POINTER(BsrStruct), P, P, c_int32, P, i32p, P], c_int),
        "gbo_pagerank": ([POINTER(BsrStruct), P, c_float, c_double, c_int32, P, i32p, POINTER(c_double), P], c_int),
        "gbo_cc": ([POINTER(BsrStruct), P, i32p, P], c_int),
        "gbo_tc": ([c_int, c_int32, P, P, POINTER(c_uint64), P], c_int),
        "gbo_bfs_multisource": ([POINTER(BsrStruct), P, c_int32, P, i32p, P], c_int),
        "gbo_autotune": ([P, P, c_int32, c_int32, c_int, c_int, c_int, POINTER(Tuning), P], c_int),
        "gbo_batch_create": ([c_int, c_int32, P, P, P, POINTER(P), P], c_int),
        "gbo_batch_offsets": ([P, P], c_int),
        "gbo_batch_info": ([P, i32p, i64p, i64p], c_int),
        "gbo_batch_bfs_cc": ([P, P, P, P, c_int, P], c_int),
        "gbo_batch_free": ([P], None),
        "gbo_ppr_forward": ([POINTER(BsrStruct), P, c_int32, c_float, P, P, c_double, c_int32, i32p,
                             POINTER(c_double), P], c_int),
        "gbo_ppr_adjoint": ([POINTER(BsrStruct), POINTER(BsrStruct), P, c_int32, c_float, P, P, P, P,
                             POINTER(c_double), c_double, c_int32, i32p, P], c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    return L


def lib():
    """The loaded libgbo.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C graph-bit-ops_amd/csrc)")
        _LIB = _declare(ctypes.CDLL(LIB_PATH))
    return _LIB


def _check(rc: int):
    if rc != GBO_OK:
        raise GboError(rc, lib().gbo_last_error().decode(errors="replace"))


def _ptr(a):
    """Device pointer of a DeviceBuffer, host pointer of a numpy array, or None."""
    if a is None:
        return None
    if isinstance(a, DeviceBuffer):
        return a.ptr
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    if isinstance(a, int):
        return a
    raise TypeError(type(a))


def hip_runtimes() -> list[str]:
    """Distinct HIP runtime libraries (libamdhip64) mapped into this process.
    Device pointers can pass between libgbo.so and another HIP user (torch)
    only when there is one: torch wheels bundle their own copy, and the
    dynamic loader reuses it for libgbo.so (same soname) only when torch was
    loaded first."""
    found = set()
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1] if "/" in line else ""
            if os.path.basename(path).startswith("libamdhip64.so"):
                found.add(os.path.realpath(path))
    return sorted(found)


def device_count() -> int:
    n = c_int(0)
    _check(lib().gbo_device_count(byref(n)))
    return n.value


def synchronize():
    _check(lib().gbo_synchronize())


class DeviceBuffer:
    """An owned device allocation (hipMalloc through the library)."""

    def __init__(self, nbytes: int):
        p = c_void_p()
        _check(lib().gbo_malloc(byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            _check(lib().gbo_memcpy_h2d(buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceBuffer":
        buf = cls(nbytes)
        _check(lib().gbo_memset(buf.ptr, 0, max(int(nbytes), 1)))
        return buf

    def to_numpy(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(lib().gbo_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def copy_from(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes:
            _check(lib().gbo_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def view(self, offset: int, nbytes: int) -> "DeviceView":
        """A non-owning window [offset, offset + nbytes) of this allocation."""
        if offset < 0 or offset + nbytes > self.nbytes:
            raise ValueError(f"view [{offset}, {offset + nbytes}) outside {self.nbytes} bytes")
        return DeviceView(self.ptr + offset, nbytes)

    def free(self):
        if getattr(self, "ptr", None):
            lib().gbo_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ----------------------------------------------------------------- CSR host
class DeviceView(DeviceBuffer):
    """A window of a DeviceBuffer (freeing it frees nothing)."""

    def __init__(self, ptr: int, nbytes: int):  # noqa: D107 -- no allocation
        self.ptr = int(ptr)
        self.nbytes = int(nbytes)

    def free(self):
        self.ptr = None

    def __del__(self):
        pass


class Csr:
    """Host CSR (int32 rowptr/colind), as produced by the readers/generators."""

    def __init__(self, nrows: int, ncols: int, rowptr: np.ndarray, colind: np.ndarray, values=None):
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colind = np.ascontiguousarray(colind, dtype=np.int32)
        self.values = values

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def outdeg(self) -> np.ndarray:
        return np.diff(self.rowptr).astype(np.int32)

    def transpose(self) -> "Csr":
        tp, ti = c_void_p(), c_void_p()
        _check(lib().gbo_csr_transpose(self.nrows, self.ncols, _ptr(self.rowptr), _ptr(self.colind), byref(tp), byref(ti)))
        return Csr(self.ncols, self.nrows, *_take(tp, self.ncols + 1, ti, self.nnz))

    def tril(self, strict: bool = True) -> "Csr":
        nnz = c_int64()
        lp, li = c_void_p(), c_void_p()
        _check(lib().gbo_csr_tril(self.nrows, _ptr(self.rowptr), _ptr(self.colind), int(strict), byref(nnz), byref(lp), byref(li)))
        return Csr(self.nrows, self.ncols, *_take(lp, self.nrows + 1, li, nnz.value))


def _take(p_rowptr, nrp, p_colind, nnz):
    """Copy library-allocated host arrays into numpy and release them."""
    L = lib()
    rp = np.ctypeslib.as_array(ctypes.cast(p_rowptr, POINTER(c_int32)), shape=(nrp,)).copy()
    ci = (np.ctypeslib.as_array(ctypes.cast(p_colind, POINTER(c_int32)), shape=(nnz,)).copy()
          if nnz > 0 else np.zeros(0, np.int32))
    L.gbo_free_host(p_rowptr)
    L.gbo_free_host(p_colind)
    return rp, ci


def mtx_read(path: str, semantics: int = 1, directed: int = 0, values: bool = False,
             cache: bool = False) -> Csr:
    """Read a Matrix Market file. semantics 0 = mmio_highlevel.h, 1 = readMtx.hpp.
    cache=True goes through gbo_mtx_read_cached (binary CSR cache beside the
    file); the returned Csr then carries `cache_state` (1 hit, 0 written,
    2 not writable)."""
    nr, nc, nnz = c_int32(), c_int32(), c_int64()
    rp, ci, val = c_void_p(), c_void_p(), c_void_p()
    state = c_int(-1)
    if cache:
        _check(lib().gbo_mtx_read_cached(os.fsencode(path), semantics, directed, byref(state), byref(nr), byref(nc),
                                         byref(nnz), byref(rp), byref(ci), byref(val) if values else None))
    else:
        _check(lib().gbo_mtx_read(os.fsencode(path), semantics, directed, byref(nr), byref(nc), byref(nnz),
                                  byref(rp), byref(ci), byref(val) if values else None))
    vals = None
    if values:
        vals = (np.ctypeslib.as_array(ctypes.cast(val, POINTER(c_float)), shape=(nnz.value,)).copy()
                if nnz.value > 0 else np.zeros(0, np.float32))
        lib().gbo_free_host(val)
    rpa, cia = _take(rp, nr.value + 1, ci, nnz.value)
    out = Csr(nr.value, nc.value, rpa, cia, vals)
    if cache:
        out.cache_state = state.value
    return out


def mtx_cache_path(path: str, semantics: int = 1, directed: int = 0) -> str:
    """The binary cache file name mtx_read(cache=True) uses."""
    buf = ctypes.create_string_buffer(4096)
    _check(lib().gbo_mtx_cache_path(os.fsencode(path), semantics, directed, buf, len(buf)))
    return os.fsdecode(buf.value)


def mtx_write(path: str, csr: Csr):
    _check(lib().gbo_mtx_write(os.fsencode(path), csr.nrows, csr.ncols, _ptr(csr.rowptr), _ptr(csr.colind)))


def gen_er(n: int, p: float, seed: int) -> Csr:
    nnz = c_int64()
    rp, ci = c_void_p(), c_void_p()
    _check(lib().gbo_gen_er(n, p, seed, byref(nnz), byref(rp), byref(ci)))
    return Csr(n, n, *_take(rp, n + 1, ci, nnz.value))


def gen_er_m(n: int, m: int, seed: int) -> Csr:
    nnz = c_int64()
    rp, ci = c_void_p(), c_void_p()
    _check(lib().gbo_gen_er_m(n, m, seed, byref(nnz), byref(rp), byref(ci)))
    return Csr(n, n, *_take(rp, n + 1, ci, nnz.value))


def gen_rmat(scale: int, edgefactor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 1) -> Csr:
    nnz = c_int64()
    rp, ci = c_void_p(), c_void_p()
    _check(lib().gbo_gen_rmat(scale, edgefactor, a, b, c, seed, byref(nnz), byref(rp), byref(ci)))
    n = 1 << scale
    return Csr(n, n, *_take(rp, n + 1, ci, nnz.value))


def gen_king(m: int) -> Csr:
    nnz = c_int64()
    rp, ci = c_void_p(), c_void_p()
    _check(lib().gbo_gen_king(m, byref(nnz), byref(rp), byref(ci)))
    return Csr(m * m, m * m, *_take(rp, m * m + 1, ci, nnz.value))


def edge_weights(csr: Csr, lo: int = 1, hi: int = 64, seed: int = 1) -> np.ndarray:
    w = np.empty(csr.nnz, np.uint32)
    _check(lib().gbo_gen_edge_weights(csr.nrows, _ptr(csr.rowptr), _ptr(csr.colind), lo, hi, seed, _ptr(w)))
    return w


# ----------------------------------------------------------------- bit-BSR
class BsrMatrix:
    """A device-resident bit-BSR matrix (owns its gbo_bsr allocation)."""

    def __init__(self, s: BsrStruct):
        self.s = s

    @classmethod
    def from_csr(cls, csr: Csr, bs: int, direction: int = 0, plan: bool = True) -> "BsrMatrix":
        s = BsrStruct()
        _check(lib().gbo_bsr_create(bs, direction, _ptr(csr.rowptr), _ptr(csr.colind), csr.nrows, csr.ncols, byref(s), None))
        m = cls(s)
        if plan:
            m.plan()
        return m

    @classmethod
    def from_device_csr(cls, d_rowptr: DeviceBuffer, d_colind: DeviceBuffer, nrows: int, ncols: int, bs: int,
                        direction: int = 0, plan: bool = True) -> "BsrMatrix":
        s = BsrStruct()
        _check(lib().gbo_bsr_create_dev(bs, direction, d_rowptr.ptr, d_colind.ptr, nrows, ncols, byref(s), None))
        m = cls(s)
        if plan:
            m.plan()
        return m

    def plan(self, chunk_tiles: int = 0, variant: int | None = None):
        v = self.s.plan_variant if variant is None else variant
        _check(lib().gbo_bsr_plan_ex(byref(self.s), chunk_tiles, v, None))
        return self

    @property
    def ref(self):
        return byref(self.s)

    bs = property(lambda self: self.s.bs)
    nrows = property(lambda self: self.s.nrows)
    ncols = property(lambda self: self.s.ncols)
    nblockrows = property(lambda self: self.s.nblockrows)
    nblockcols = property(lambda self: self.s.nblockcols)
    nblocks = property(lambda self: self.s.nblocks)

    def tile_bytes(self) -> int:
        return self.s.nblocks * self.s.bs * self.s.bs // 8

    def host_arrays(self):
        """(rowptr, colind, tiles) copied to the host."""
        L = lib()
        s = self.s
        rp = np.empty(s.nblockrows + 1, np.int32)
        ci = np.empty(s.nblocks, np.int32)
        tw = np.empty(s.nblocks * s.bs, WORD_DTYPE[s.bs])
        _check(L.gbo_memcpy_d2h(rp.ctypes.data, s.rowptr, rp.nbytes))
        if s.nblocks:
            _check(L.gbo_memcpy_d2h(ci.ctypes.data, s.colind, ci.nbytes))
            _check(L.gbo_memcpy_d2h(tw.ctypes.data, s.tiles, tw.nbytes))
        return rp, ci, tw

    def free(self):
        if getattr(self, "s", None) is not None:
            lib().gbo_bsr_free(byref(self.s))
            self.s = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bsrbmm(A: BsrMatrix, B: BsrMatrix, counts: bool = False):
    """General bit-BSR SpGEMM C = A * B (both row-packed): the Boolean product
    as a library-owned BsrMatrix; with counts=True also (valoff, counts) as
    host arrays -- the integer product at C's set bits in tile bit order."""
    s = BsrStruct()
    _check(lib().gbo_bsrbmm(A.ref, B.ref, byref(s), None))
    C = BsrMatrix(s)
    if not counts:
        return C
    nb = C.nblocks
    d_off = DeviceBuffer(8 * (nb + 1))
    nnz = c_int64()
    _check(lib().gbo_bsr_bit_offsets(C.ref, d_off.ptr, byref(nnz), None))
    d_cnt = DeviceBuffer(4 * max(nnz.value, 1))
    _check(lib().gbo_bsrbmm_values(A.ref, B.ref, C.ref, d_off.ptr, d_cnt.ptr, None))
    off = d_off.to_numpy(np.int64, nb + 1)
    cnt = d_cnt.to_numpy(np.int32, nnz.value)
    d_off.free()
    d_cnt.free()
    return C, off, cnt


def bsrbmm_phases(A: BsrMatrix, B: BsrMatrix):
    """The two-phase, caller-allocates form (symbolic sizes C, numeric fills
    it), as the reference's bmm32_sparse_symbolic / _numeric pair is called.
    Returns host (rowptr, colind, tiles, valoff, nnz)."""
    L = lib()
    nbr = A.nblockrows
    d_rp = DeviceBuffer(4 * (nbr + 1))
    nb = c_int64()
    _check(L.gbo_bsrbmm_symbolic(A.ref, B.ref, d_rp.ptr, byref(nb), None))
    n = nb.value
    bs = A.bs
    d_ci = DeviceBuffer(4 * max(n, 1))
    d_t = DeviceBuffer(max(n * bs * bs // 8, 8))
    d_off = DeviceBuffer(8 * (n + 1))
    nnz = c_int64()
    _check(L.gbo_bsrbmm_numeric(A.ref, B.ref, d_rp.ptr, n, d_ci.ptr, d_t.ptr, d_off.ptr, byref(nnz), None))
    out = (d_rp.to_numpy(np.int32, nbr + 1), d_ci.to_numpy(np.int32, n),
           d_t.to_numpy(WORD_DTYPE[bs], n * bs), d_off.to_numpy(np.int64, n + 1), nnz.value)
    for b in (d_rp, d_ci, d_t, d_off):
        b.free()
    return out


def csr2bsr_batch(csr: Csr, bs: int, direction: int = 0):
    """Reference-shaped packing call: sizes with gbo_csr2bsr_nnzb (the
    cusparseXcsr2bsrNnz step), allocates d_bsrRowPtr/d_bsrColInd/d_bsrVal like
    the reference caller (bsrbmv/benchmark-bsrbmv.cu:1036-1048), then packs.
    Returns (d_rowptr, d_colind, d_tiles, nblockrows, nblocks)."""
    L = lib()
    nb = c_int64()
    _check(L.gbo_csr2bsr_nnzb(bs, _ptr(csr.rowptr), _ptr(csr.colind), csr.nrows, csr.ncols, byref(nb)))
    nbr = (csr.nrows + bs - 1) // bs
    d_rp = DeviceBuffer(4 * (nbr + 1))
    d_ci = DeviceBuffer(4 * max(nb.value, 1))
    d_t = DeviceBuffer(max(nb.value * bs * bs // 8, 8))
    _check(L.gbo_csr2bsr_batch(bs, direction, _ptr(csr.rowptr), _ptr(csr.colind), csr.nrows, csr.ncols, csr.nnz,
                               d_rp.ptr, d_ci.ptr, d_t.ptr, nbr, nb.value, None))
    return d_rp, d_ci, d_t, nbr, nb.value


def pack_vec(bs: int, x: np.ndarray) -> np.ndarray:
    """ToBit{N}Row on the GPU: float vector -> bit words (element > 0)."""
    x = np.ascontiguousarray(x, np.float32)
    nw = (len(x) + bs - 1) // bs
    dx = DeviceBuffer.from_numpy(x)
    dw = DeviceBuffer(max(nw, 1) * bs // 8)
    _check(lib().gbo_pack_vec(bs, dx.ptr, dw.ptr, len(x), None))
    return dw.to_numpy(WORD_DTYPE[bs], nw)


def _dev(a, dtype=None) -> DeviceBuffer:
    if isinstance(a, DeviceBuffer):
        return a
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    return DeviceBuffer.from_numpy(a)


def bsrbmv_count(A: BsrMatrix, xbits, y_init: np.ndarray | None = None) -> np.ndarray:
    """int32 y = A * x over bit words; with y_init the product is added (+=)."""
    dx = _dev(xbits, WORD_DTYPE[A.bs])
    if y_init is not None:
        dy = DeviceBuffer.from_numpy(np.asarray(y_init, np.int32))
    else:
        dy = DeviceBuffer(4 * max(A.nrows, 1))
    _check(lib().gbo_bsrbmv_count(A.ref, dx.ptr, dy.ptr, int(y_init is not None), None))
    return dy.to_numpy(np.int32, A.nrows)


def bsrbmv_bfs_step(A: BsrMatrix, frontier, visited):
    """(next bit words, popcount) of one masked pull-BFS level."""
    W = WORD_DTYPE[A.bs]
    df, dv = _dev(frontier, W), _dev(visited, W)
    dn = DeviceBuffer(max(A.nblockrows, 1) * A.bs // 8)
    dc = DeviceBuffer(8)
    _check(lib().gbo_bsrbmv_bfs_step(A.ref, df.ptr, dv.ptr, dn.ptr, dc.ptr, None))
    return dn.to_numpy(W, A.nblockrows), int(dc.to_numpy(np.int64, 1)[0])


class TileValues:
    """Per-nonzero uint32 values in tile bit order (+ per-tile offsets)."""

    def __init__(self, A: BsrMatrix, csr: Csr, values: np.ndarray):
        d_rp = DeviceBuffer.from_numpy(csr.rowptr)
        d_ci = DeviceBuffer.from_numpy(csr.colind)
        d_v = DeviceBuffer.from_numpy(np.asarray(values, np.uint32))
        self.tval = DeviceBuffer(4 * max(csr.nnz, 1))
        self.woff = DeviceBuffer(8 * (A.nblocks + 1))
        _check(lib().gbo_csr2bsr_values_u32(A.ref, d_rp.ptr, d_ci.ptr, d_v.ptr, self.tval.ptr, self.woff.ptr, None))
        synchronize()
        self.count = int(self.woff.to_numpy(np.int64)[A.nblocks]) if A.nblocks else 0

    def free(self):
        self.tval.free()
        self.woff.free()


def bsrbmv_minplus(A: BsrMatrix, dist, weights: TileValues | None = None) -> np.ndarray:
    dd = _dev(dist, np.uint32)
    do = DeviceBuffer(4 * max(A.nrows, 1))
    w = weights.tval.ptr if weights else None
    wo = weights.woff.ptr if weights else None
    _check(lib().gbo_bsrbmv_minplus(A.ref, w, wo, dd.ptr, do.ptr, None))
    return do.to_numpy(np.uint32, A.nrows)


def bsrbmv_sum_f32(A: BsrMatrix, x) -> np.ndarray:
    dx = _dev(x, np.float32)
    dy = DeviceBuffer(4 * max(A.nrows, 1))
    _check(lib().gbo_bsrbmv_sum_f32(A.ref, dx.ptr, dy.ptr, None))
    return dy.to_numpy(np.float32, A.nrows)


def bsrbmv_min_i32(A: BsrMatrix, x) -> np.ndarray:
    dx = _dev(x, np.int32)
    dy = DeviceBuffer(4 * max(A.nrows, 1))
    _check(lib().gbo_bsrbmv_min_i32(A.ref, dx.ptr, dy.ptr, None))
    return dy.to_numpy(np.int32, A.nrows)


def bsrbmm_tc(L_row: BsrMatrix, LT_col: BsrMatrix) -> int:
    d = DeviceBuffer(8)
    _check(lib().gbo_bsrbmm_tc(L_row.ref, LT_col.ref, d.ptr, None))
    return int(d.to_numpy(np.uint64, 1)[0])


def bsrbmm_sum(A_row: BsrMatrix, B_col: BsrMatrix) -> int:
    d = DeviceBuffer(8)
    _check(lib().gbo_bsrbmm_sum(A_row.ref, B_col.ref, d.ptr, None))
    synchronize()
    return int(d.to_numpy(np.uint64, 1)[0])


# ----------------------------------------------------------------- drivers
def _result(dbuf: DeviceBuffer, dtype, n: int, out: DeviceBuffer | None):
    """The result as a host array, or the caller's device buffer as it is."""
    return out if out is not None else dbuf.to_numpy(dtype, n)


def _out_buffer(out: DeviceBuffer | None, nbytes: int) -> DeviceBuffer:
    if out is None:
        return DeviceBuffer(max(nbytes, 4))
    if out.nbytes < nbytes:
        raise ValueError(f"output buffer of {out.nbytes} bytes, {nbytes} needed")
    return out


def bfs(AT: BsrMatrix, source: int = 0, max_levels: int = 4096, A: BsrMatrix | None = None,
        out: DeviceBuffer | None = None):
    """(levels int32[n], per-level discovered counts).  With A (the out-edge
    matrix; AT itself for an undirected graph) small frontiers are pushed
    (gbo_bfs_do); without it every level is pulled over AT (gbo_bfs).  With
    out (a device buffer of >= 4n bytes) the levels stay there and out is
    returned in place of the host array."""
    dl = _out_buffer(out, 4 * AT.nrows)
    sizes = np.zeros(max_levels, np.int64)
    nl = c_int32()
    if A is None:
        _check(lib().gbo_bfs(AT.ref, source, dl.ptr, _ptr(sizes), max_levels, byref(nl), None))
    else:
        _check(lib().gbo_bfs_do(AT.ref, A.ref, source, dl.ptr, _ptr(sizes), max_levels, byref(nl), None))
    return _result(dl, np.int32, AT.nrows, out), sizes[: min(nl.value, max_levels)].copy()


def sssp(AT: BsrMatrix, source: int = 0, weights: TileValues | None = None, out: DeviceBuffer | None = None):
    """(uint32 distances, iterations); INF = 0xFFFFFFFF.  out as for bfs."""
    dd = _out_buffer(out, 4 * AT.nrows)
    it = c_int32()
    w = weights.tval.ptr if weights else None
    wo = weights.woff.ptr if weights else None
    _check(lib().gbo_sssp(AT.ref, w, wo, source, dd.ptr, byref(it), None))
    return _result(dd, np.uint32, AT.nrows, out), it.value


def pagerank(AT: BsrMatrix, outdeg, alpha: float = 0.85, tol: float = 1e-7, max_iter: int = 1000,
             strict: bool = False, out: DeviceBuffer | None = None):
    """(float32 scores, iterations, last L1 change).  Reaching max_iter first
    (GBO_ENOCONV) raises only when strict=True; the iterate is returned.
    outdeg is a host array or a device buffer of int32; out as for bfs."""
    do = _dev(outdeg, np.int32)
    dp = _out_buffer(out, 4 * AT.nrows)
    it = c_int32()
    delta = c_double()
    rc = lib().gbo_pagerank(AT.ref, do.ptr, alpha, tol, max_iter, dp.ptr, byref(it), byref(delta), None)
    if rc != GBO_OK and (strict or rc != -7):
        _check(rc)
    return _result(dp, np.float32, AT.nrows, out), it.value, delta.value


def cc(A: BsrMatrix, out: DeviceBuffer | None = None):
    """(int32 labels = smallest vertex id of the component, iterations).  out
    as for bfs."""
    dl = _out_buffer(out, 4 * A.nrows)
    it = c_int32()
    _check(lib().gbo_cc(A.ref, dl.ptr, byref(it), None))
    return _result(dl, np.int32, A.nrows, out), it.value


def tc(csr: Csr, bs: int) -> int:
    out = c_uint64()
    _check(lib().gbo_tc(bs, csr.nrows, _ptr(csr.rowptr), _ptr(csr.colind), byref(out), None))
    return out.value


# ----------------------------------------------------------------- batching
def bfs_multisource(AT: BsrMatrix, sources, out: DeviceBuffer | None = None):
    """Levels of up to 256 BFS runs in one batched pass: (int32 [nsrc][n], nlevels).
    With `out` given (a device buffer of nsrc*n int32) the levels stay on the
    device and only nlevels is returned."""
    src = np.ascontiguousarray(sources, np.int32)
    n = AT.nrows
    d = _out_buffer(out, 4 * len(src) * n)
    nl = c_int32()
    _check(lib().gbo_bfs_multisource(AT.ref, _ptr(src), len(src), d.ptr, byref(nl), None))
    if out is not None:
        return nl.value
    return d.to_numpy(np.int32, len(src) * n).reshape(len(src), n), nl.value


# ------------------------------------------------------ personalized PageRank
def ppr_forward(AT: BsrMatrix, outdeg, S, alpha: float = 0.85, tol: float = 1e-7, max_iter: int = 1000,
                strict: bool = False):
    """P[n][K] = personalized PageRank of the K columns of S[n][K] (float32).
    Returns (P, iterations, last max-column L1 change relative to ||s_k||_1)."""
    S = np.ascontiguousarray(S, np.float32)
    n, K = S.shape
    do = DeviceBuffer.from_numpy(np.asarray(outdeg, np.int32))
    ds = DeviceBuffer.from_numpy(S)
    dp = DeviceBuffer(4 * max(n * K, 1))
    it, delta = c_int32(), c_double()
    rc = lib().gbo_ppr_forward(AT.ref, do.ptr, K, alpha, ds.ptr, dp.ptr, tol, max_iter, byref(it), byref(delta), None)
    if rc != GBO_OK and (strict or rc != -7):
        _check(rc)
    return dp.to_numpy(np.float32, n * K).reshape(n, K), it.value, delta.value


def ppr_forward_ptr(AT: BsrMatrix, d_outdeg: int, K: int, alpha: float, d_s: int, d_p: int, tol: float = 1e-7,
                    max_iter: int = 1000, strict: bool = True, stream: int | None = None):
    """gbo_ppr_forward on caller-owned device memory (raw device addresses of
    [n][K] float32 S and P and of int32 out-degrees), stream-ordered on
    `stream`: nothing is copied to or from the host.  Returns (iterations,
    last relative change)."""
    it, delta = c_int32(), c_double()
    rc = lib().gbo_ppr_forward(AT.ref, d_outdeg, K, alpha, d_s, d_p, tol, max_iter, byref(it), byref(delta), stream)
    if rc != GBO_OK and (strict or rc != -7):
        _check(rc)
    return it.value, delta.value


def ppr_adjoint_ptr(A: BsrMatrix, AT: BsrMatrix, d_outdeg: int, K: int, alpha: float, d_s: int, d_p: int, d_w: int,
                    d_grad_s: int, tol: float = 1e-7, max_iter: int = 1000, strict: bool = True,
                    stream: int | None = None):
    """gbo_ppr_adjoint on caller-owned device memory (as ppr_forward_ptr);
    grad_S goes to d_grad_s.  Returns (grad_alpha, iterations)."""
    ga, it = c_double(), c_int32()
    rc = lib().gbo_ppr_adjoint(A.ref, AT.ref, d_outdeg, K, alpha, d_s, d_p, d_w, d_grad_s, byref(ga), tol, max_iter,
                               byref(it), stream)
    if rc != GBO_OK and (strict or rc != -7):
        _check(rc)
    return ga.value, it.value


def ppr_adjoint(A: BsrMatrix, AT: BsrMatrix, outdeg, S, P, W, alpha: float = 0.85, tol: float = 1e-7,
                max_iter: int = 1000, strict: bool = False):
    """Gradients of L = sum(W * P) with respect to S and alpha:
    (grad_S[n][K] float32, grad_alpha float, iterations).  The adjoint solve
    stops when the max-column L1 change relative to ||W_k||_1 is below tol."""
    S = np.ascontiguousarray(S, np.float32)
    n, K = S.shape
    do = DeviceBuffer.from_numpy(np.asarray(outdeg, np.int32))
    ds = DeviceBuffer.from_numpy(S)
    dp = DeviceBuffer.from_numpy(np.ascontiguousarray(P, np.float32))
    dw = DeviceBuffer.from_numpy(np.ascontiguousarray(W, np.float32))
    dg = DeviceBuffer(4 * max(n * K, 1))
    ga, it = c_double(), c_int32()
    rc = lib().gbo_ppr_adjoint(A.ref, AT.ref, do.ptr, K, alpha, ds.ptr, dp.ptr, dw.ptr, dg.ptr, byref(ga), tol,
                               max_iter, byref(it), None)
    if rc != GBO_OK and (strict or rc != -7):
        _check(rc)
    return dg.to_numpy(np.float32, n * K).reshape(n, K), ga.value, it.value


# ------------------------------------------------------------ graph batch
class GraphBatch:
    """Many graphs packed side by side; BFS + CC of all of them in one
    persistent (optionally hipGraph-captured) launch."""

    def __init__(self, graphs, bs: int):
        L = lib()
        G = len(graphs)
        self.graphs = graphs
        nrows = np.array([g.nrows for g in graphs], np.int32)
        rps = (c_void_p * G)(*[g.rowptr.ctypes.data for g in graphs])
        cis = (c_void_p * G)(*[g.colind.ctypes.data if g.nnz else 0 for g in graphs])
        h = c_void_p()
        _check(L.gbo_batch_create(bs, G, nrows.ctypes.data, ctypes.cast(rps, c_void_p), ctypes.cast(cis, c_void_p),
                                  byref(h), None))
        self.h = h.value
        self.offsets = np.zeros(G + 1, np.int64)
        _check(L.gbo_batch_offsets(self.h, self.offsets.ctypes.data))
        bsv, nb, nv = c_int32(), c_int64(), c_int64()
        _check(L.gbo_batch_info(self.h, byref(bsv), byref(nb), byref(nv)))
        self.bs, self.nblocks, self.nvertices_padded = bsv.value, nb.value, nv.value
        total = int(self.offsets[-1])
        self.d_level = DeviceBuffer(4 * max(total, 1))
        self.d_label = DeviceBuffer(4 * max(total, 1))

    def run(self, sources, use_graph: bool = True):
        src = np.ascontiguousarray(sources, np.int32)
        _check(lib().gbo_batch_bfs_cc(self.h, src.ctypes.data, self.d_level.ptr, self.d_label.ptr, int(use_graph), None))

    def results(self):
        total = int(self.offsets[-1])
        lv = self.d_level.to_numpy(np.int32, total)
        lb = self.d_label.to_numpy(np.int32, total)
        o = self.offsets
        return [lv[o[i]:o[i + 1]] for i in range(len(o) - 1)], [lb[o[i]:o[i + 1]] for i in range(len(o) - 1)]

    def free(self):
        if getattr(self, "h", None):
            lib().gbo_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gen_er_batch(count: int = 512, n_lo: int = 1000, n_hi: int = 65536, d_lo: float = 8.0, d_hi: float = 32.0,
                 seed: int = 4):
    """Erdos-Renyi graphs with n log-uniform in [n_lo, n_hi] and mean degree
    uniform in [d_lo, d_hi] (G(n, m) form, m = n * d / 2 draws)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        d = rng.uniform(d_lo, d_hi)
        out.append(gen_er_m(n, int(n * d / 2), seed * 1_000_003 + i))
    return out


def autotune(d_rowptr: DeviceBuffer, d_colind: DeviceBuffer, nrows: int, ncols: int, op: int = 0,
             block_sizes=(8, 16, 32, 64), reps: int = 3) -> Tuning:
    """Fastest (bs, chunk, variant) for a bsrbmv sweep of this device CSR."""
    mask = sum(1 << {8: 0, 16: 1, 32: 2, 64: 3}[b] for b in block_sizes)
    t = Tuning()
    _check(lib().gbo_autotune(d_rowptr.ptr, d_colind.ptr, nrows, ncols, op, mask, reps, byref(t), None))
    return t
