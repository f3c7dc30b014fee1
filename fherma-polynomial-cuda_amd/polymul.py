"""Python binding of the gfx950 negacyclic multiplier (ctypes over the C ABI of
include/polymul.h, libpolymul_amd.so in this directory).

The library is the product; this module only marshals numpy arrays and raw
device pointers.  It never computes a product itself: if the shared library is
missing, loading fails with an error telling how to build it.
"""
import ctypes
import threading
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY_PATH = HERE / "libpolymul_amd.so"

PM_OK = 0
PM_ERR_INVALID_ARGUMENT = 1
PM_ERR_INSUFFICIENT_FIELDS = 2
PM_ERR_DEVICE = 3
PM_ERR_OUT_OF_MEMORY = 4
PM_ERR_INTERNAL = 5
PM_ERR_NO_DEVICE = 6
STATUS_NAMES = {
    0: "PM_OK",
    1: "PM_ERR_INVALID_ARGUMENT",
    2: "PM_ERR_INSUFFICIENT_FIELDS",
    3: "PM_ERR_DEVICE",
    4: "PM_ERR_OUT_OF_MEMORY",
    5: "PM_ERR_INTERNAL",
    6: "PM_ERR_NO_DEVICE",
}

TABLE_PRIME_CONSTS = 1
TABLE_CONV_WEIGHTS = 2
TABLE_TWIDDLES_FWD = 3
TABLE_TWIDDLES_INV = 4
TABLE_CRT_WEIGHTS = 5
TABLE_M_MOD_Q = 6
TABLE_NEG_P_MOD_Q = 7
TABLE_CORR_T = 8
TABLE_CORR_E = 9
TABLE_RECIPROCALS = 10
TABLE_BETA = 11
TABLE_SHAPE = 12

SCRATCH_BYTES = 128 << 20  # PM_SCRATCH_BYTES: static workspace of pm_multiply_device_scratch

STAGE_RESIDUES = 1
STAGE_PRODUCT = 2
STAGE_WORDS = 3
STAGE_OUTPUT = 4

SHAPE_FIELDS = ("n", "width", "limbs", "n1", "n2", "log_n2", "r", "r_pad", "conv_k", "conv_rows",
                "crt_words", "crt_words_pad", "crt_rows", "crt_k", "t_words", "row_width", "rows",
                "qbits", "s_shift", "k_shift", "mu_lo", "mu_hi", "trivial", "reference_walk")

STAGE_TIME_NAMES = ("pack_operands", "conversion_gemm", "ntt_columns_forward", "ntt_rows_product",
                    "ntt_columns_inverse", "crt_gemm", "reconstruct", "total")


class PolymulError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status


class PlanInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("width", ctypes.c_uint32), ("limbs", ctypes.c_uint32),
                ("primes", ctypes.c_uint32), ("block", ctypes.c_uint32), ("transform", ctypes.c_uint32),
                ("row_width", ctypes.c_uint32), ("reference_walk", ctypes.c_uint32),
                ("qbits", ctypes.c_uint32), ("trivial", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64)]


class Request(ctypes.Structure):
    """pm_request: one product with its own point and host operands."""
    _fields_ = [("n", ctypes.c_uint32), ("width", ctypes.c_uint32), ("limbs", ctypes.c_uint32),
                ("q", ctypes.c_void_p), ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p)]


class RequestTimes(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("service_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


_lib = None


def hip_runtimes():
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        maps = Path("/proc/self/maps").read_text()
    except OSError:
        return set()
    return {line.split()[-1] for line in maps.splitlines() if "libamdhip64.so" in line.split()[-1]}


def _bind_one_hip_runtime():
    """The library links ROCm's HIP runtime by its SONAME (libamdhip64.so.7);
    PyTorch's ROCm wheel ships its own copy and links it by the unversioned
    name.  Opening the runtime under the unversioned name first makes both
    resolve to one copy, whichever of the two is loaded first: if torch is
    already loaded this returns its copy, otherwise ROCm's copy is registered
    under both names and a later `import torch` reuses it."""
    try:
        ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
    except OSError:
        pass  # the library's own NEEDED entry still finds a runtime


def load_library(path=None):
    """Loads libpolymul_amd.so; raises (never falls back) when it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = Path(path) if path else LIBRARY_PATH
    if not path.is_file():
        raise FileNotFoundError("%s is not built; run `python fherma-polynomial-cuda_amd/build.py`" % path)
    _bind_one_hip_runtime()
    lib = ctypes.CDLL(str(path))
    if len(hip_runtimes()) > 1:
        raise RuntimeError("two HIP runtimes are loaded in this process (%s); the second sees no GPU"
                           % ", ".join(sorted(hip_runtimes())))
    u32p = ctypes.POINTER(ctypes.c_uint32)
    vp = ctypes.c_void_p
    plan_pp = ctypes.POINTER(ctypes.c_void_p)
    sigs = {
        "pm_plan_create": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p, plan_pp], ctypes.c_int),
        "pm_plan_create_host": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p, plan_pp], ctypes.c_int),
        "pm_plan_create_split": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32,
                                  ctypes.c_int, plan_pp], ctypes.c_int),
        "pm_plan_destroy": ([vp], None),
        "pm_multiply": ([vp, vp, vp, vp], ctypes.c_int),
        "pm_multiply_device": ([vp, vp, vp, vp, vp], ctypes.c_int),
        "pm_multiply_device_scratch": ([vp, vp, vp, vp, vp], ctypes.c_int),
        "pm_multiply_device_batch": ([vp, vp, vp, vp, ctypes.c_size_t, vp], ctypes.c_int),
        "pm_multiply_begin": ([vp, vp, vp], ctypes.c_int),
        "pm_multiply_finish": ([vp, vp], ctypes.c_int),
        "pm_last_error": ([], ctypes.c_char_p),
        "pm_plan_get_info": ([vp, ctypes.POINTER(PlanInfo)], ctypes.c_int),
        "pm_plan_get_primes": ([vp, u32p, ctypes.c_size_t], ctypes.c_int),
        "pm_plan_export_table": ([vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "pm_debug_stage": ([vp, vp, vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "pm_set_profiling": ([vp, ctypes.c_int], ctypes.c_int),
        "pm_operand_create": ([vp, vp, plan_pp], ctypes.c_int),
        "pm_multiply_operand": ([vp, vp, vp, vp], ctypes.c_int),
        "pm_multiply_operand_device": ([vp, vp, vp, vp, vp], ctypes.c_int),
        "pm_operand_destroy": ([vp], None),
        "pm_get_stage_times": ([vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int], ctypes.c_int),
        "pm_server_create": ([ctypes.c_uint64, ctypes.c_uint32, plan_pp], ctypes.c_int),
        "pm_server_submit": ([vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "pm_server_wait": ([vp, ctypes.c_uint64, ctypes.POINTER(RequestTimes)], ctypes.c_int),
        "pm_server_destroy": ([vp], None),
        "pm_multiply_batch": ([ctypes.POINTER(Request), ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint32,
                               ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path == LIBRARY_PATH:
        _lib = lib
    return lib


def _check(lib, status):
    if status != PM_OK:
        raise PolymulError(status, lib.pm_last_error().decode(errors="replace"))


def _u32(array):
    array = np.ascontiguousarray(array, dtype=np.uint32)
    return array, array.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class Plan:
    """A parameter-only plan (fherma_init) for c = a*b mod (X^N+1, q)."""

    def __init__(self, n, width, limbs, q_words, device=True, min_block=None):
        """min_block: force the transform split to start at n1 >= min_block
        (pm_plan_create_split); None takes the plain choice."""
        self.lib = load_library()
        q_arr, q_ptr = _u32(q_words)
        if q_arr.size != limbs:
            raise ValueError("q must have exactly L=%d limbs" % limbs)
        handle = ctypes.c_void_p()
        if min_block is not None:
            status = self.lib.pm_plan_create_split(n, width, limbs, q_ptr, min_block, int(device), ctypes.byref(handle))
        else:
            create = self.lib.pm_plan_create if device else self.lib.pm_plan_create_host
            status = create(n, width, limbs, q_ptr, ctypes.byref(handle))
        _check(self.lib, status)
        self.handle = handle
        self.n, self.width, self.limbs = n, width, limbs
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self.lib.pm_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- inspection -----------------------------------------------------
    def info(self):
        info = PlanInfo()
        _check(self.lib, self.lib.pm_plan_get_info(self.handle, ctypes.byref(info)))
        return {name: getattr(info, name) for name, _ in PlanInfo._fields_}

    def primes(self):
        count = self.info()["primes"]
        out = np.zeros(count, dtype=np.uint32)
        _check(self.lib, self.lib.pm_plan_get_primes(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                     count))
        return out

    def table(self, table_id, dtype):
        size = ctypes.c_size_t(0)
        _check(self.lib, self.lib.pm_plan_export_table(self.handle, table_id, None, ctypes.byref(size)))
        out = np.zeros(size.value // np.dtype(dtype).itemsize, dtype=dtype)
        _check(self.lib, self.lib.pm_plan_export_table(self.handle, table_id, out.ctypes.data_as(ctypes.c_void_p),
                                                       ctypes.byref(size)))
        return out

    def shape(self):
        words = self.table(TABLE_SHAPE, np.uint32)
        s = dict(zip(SHAPE_FIELDS, (int(w) for w in words)))
        s["mu"] = s.pop("mu_lo") | (s.pop("mu_hi") << 32)
        return s

    # ---- compute --------------------------------------------------------
    def _out(self, out):
        """`out`, checked to be a writeable contiguous uint32 [N][L] array, or a
        new one.  A caller running many products can pass the same array each
        time: a fresh one costs its page faults on the library's first write,
        about as long as a product at N = 65536 with 55 limbs."""
        if out is None:
            return np.empty((self.n, self.limbs), dtype=np.uint32)
        if (not isinstance(out, np.ndarray) or out.shape != (self.n, self.limbs) or out.dtype != np.uint32
                or not out.flags.c_contiguous or not out.flags.writeable):
            raise ValueError("out must be a writeable contiguous uint32 [N][L] array")
        return out

    def multiply(self, a, b, out=None):
        """Host [N][L] uint32 arrays in, host array out (includes transfers);
        c is written into `out` when given."""
        a, _ = _u32(a)
        b, _ = _u32(b)
        expect = (self.n, self.limbs)
        if a.shape != expect or b.shape != expect:
            raise ValueError("operands must be [N][L] = %s" % (expect,))
        c = self._out(out)
        _check(self.lib, self.lib.pm_multiply(self.handle, a.ctypes.data_as(ctypes.c_void_p),
                                              b.ctypes.data_as(ctypes.c_void_p), c.ctypes.data_as(ctypes.c_void_p)))
        return c

    def multiply_begin(self, a, b):
        """First half of multiply(): inputs in, device work queued."""
        a, _ = _u32(a)
        b, _ = _u32(b)
        expect = (self.n, self.limbs)
        if a.shape != expect or b.shape != expect:
            raise ValueError("operands must be [N][L] = %s" % (expect,))
        _check(self.lib, self.lib.pm_multiply_begin(self.handle, a.ctypes.data_as(ctypes.c_void_p),
                                                    b.ctypes.data_as(ctypes.c_void_p)))

    def multiply_finish(self, out=None):
        """Second half: waits and returns c (written into `out` if given)."""
        c = self._out(out)
        _check(self.lib, self.lib.pm_multiply_finish(self.handle, c.ctypes.data_as(ctypes.c_void_p)))
        return c

    def multiply_device(self, a_ptr, b_ptr, c_ptr, stream=None):
        """Device pointers (ints) to [N][L] words; stream is a hipStream_t int or None."""
        _check(self.lib, self.lib.pm_multiply_device(self.handle, ctypes.c_void_p(a_ptr), ctypes.c_void_p(b_ptr),
                                                     ctypes.c_void_p(c_ptr),
                                                     ctypes.c_void_p(stream) if stream else None))

    def multiply_device_scratch(self, a_ptr, b_ptr, c_ptr, stream=None):
        """multiply_device with the working buffers in the library's static
        per-GPU workspace (pm_multiply_device_scratch): the call allocates
        nothing; plans whose workspace exceeds SCRATCH_BYTES are refused."""
        _check(self.lib, self.lib.pm_multiply_device_scratch(self.handle, ctypes.c_void_p(a_ptr),
                                                             ctypes.c_void_p(b_ptr), ctypes.c_void_p(c_ptr),
                                                             ctypes.c_void_p(stream) if stream else None))

    def debug_stage(self, a, b, stage):
        a, _ = _u32(a)
        b, _ = _u32(b)
        size = ctypes.c_size_t(0)
        _check(self.lib, self.lib.pm_debug_stage(self.handle, None, None, stage, None, ctypes.byref(size)))
        dtype = np.int64 if stage == STAGE_WORDS else np.uint32
        out = np.zeros(size.value // np.dtype(dtype).itemsize, dtype=dtype)
        _check(self.lib, self.lib.pm_debug_stage(self.handle, a.ctypes.data_as(ctypes.c_void_p),
                                                 b.ctypes.data_as(ctypes.c_void_p), stage,
                                                 out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(size)))
        return out

    # ---- resident operand -------------------------------------------------
    def operand(self, b):
        """Converts and transforms b once; returns a ResidentOperand."""
        b, _ = _u32(b)
        if b.shape != (self.n, self.limbs):
            raise ValueError("operand must be [N][L]")
        handle = ctypes.c_void_p()
        _check(self.lib, self.lib.pm_operand_create(self.handle, b.ctypes.data_as(ctypes.c_void_p),
                                                    ctypes.byref(handle)))
        return ResidentOperand(self, handle)

    def multiply_operand(self, a, operand, out=None):
        """a times a resident operand; c is written into `out` when given."""
        a, _ = _u32(a)
        if a.shape != (self.n, self.limbs):
            raise ValueError("operand must be [N][L]")
        c = self._out(out)
        _check(self.lib, self.lib.pm_multiply_operand(self.handle, a.ctypes.data_as(ctypes.c_void_p), operand.handle,
                                                      c.ctypes.data_as(ctypes.c_void_p)))
        return c

    def multiply_operand_device(self, a_ptr, operand, c_ptr, stream=None):
        _check(self.lib, self.lib.pm_multiply_operand_device(self.handle, ctypes.c_void_p(a_ptr), operand.handle,
                                                             ctypes.c_void_p(c_ptr),
                                                             ctypes.c_void_p(stream) if stream else None))

    def set_profiling(self, enabled=True):
        _check(self.lib, self.lib.pm_set_profiling(self.handle, 1 if enabled else 0))

    def stage_times(self):
        out = (ctypes.c_float * len(STAGE_TIME_NAMES))()
        _check(self.lib, self.lib.pm_get_stage_times(self.handle, out, len(STAGE_TIME_NAMES)))
        return dict(zip(STAGE_TIME_NAMES, (float(x) for x in out)))


class ResidentOperand:
    """A factor b whose transformed residues stay on the device (pm_operand)."""

    def __init__(self, plan, handle):
        self.plan = plan
        self.handle = handle

    def close(self):
        if getattr(self, "handle", None):
            self.plan.lib.pm_operand_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _request(n, limbs, q, a, b):
    """Validated contiguous buffers and a pm_request for one product; returns
    (request, c, keepalive)."""
    q, _ = _u32(q)
    a, _ = _u32(a)
    b, _ = _u32(b)
    if q.size != limbs or a.shape != (n, limbs) or b.shape != (n, limbs):
        raise ValueError("request buffers must be q[L], a[N][L], b[N][L]")
    c = np.empty((n, limbs), dtype=np.uint32)
    req = Request(n, 32 * limbs, limbs, q.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data)
    return req, c, (q, a, b)


def multiply_batch(products, device_budget_bytes, lanes=4):
    """Many products, each (n, limbs, q, a, b) with its own q, through one
    temporary server (pm_multiply_batch).  Returns the list of c arrays."""
    lib = load_library()
    built = [_request(*p) for p in products]
    reqs = (Request * len(built))(*[b[0] for b in built])
    statuses = (ctypes.c_int * len(built))()
    _check(lib, lib.pm_multiply_batch(reqs, len(built), int(device_budget_bytes), int(lanes), statuses))
    return [b[1] for b in built]


def multiply_device_batch(plans, a_ptrs, b_ptrs, c_ptrs, stream=None):
    """Many device-resident products, plans[i]: a[i] * b[i] -> c[i] (device
    pointers as ints), each stage one launch for the whole batch
    (pm_multiply_device_batch)."""
    lib = load_library()
    count = len(plans)
    if not (len(a_ptrs) == len(b_ptrs) == len(c_ptrs) == count):
        raise ValueError("one a, b and c pointer per plan")
    handles = (ctypes.c_void_p * count)(*[p.handle.value for p in plans])
    arrays = [(ctypes.c_void_p * count)(*[int(x) for x in xs]) for xs in (a_ptrs, b_ptrs, c_ptrs)]
    _check(lib, lib.pm_multiply_device_batch(handles, arrays[0], arrays[1], arrays[2], count,
                                             ctypes.c_void_p(stream) if stream else None))


class Server:
    """A request server (pm_server): submit() queues a product and returns a
    ticket at once; wait(ticket) returns (c, times) when it has finished."""

    def __init__(self, device_budget_bytes, lanes=4):
        self.lib = load_library()
        handle = ctypes.c_void_p()
        _check(self.lib, self.lib.pm_server_create(int(device_budget_bytes), int(lanes), ctypes.byref(handle)))
        self.handle = handle
        self._pending = {}  # ticket -> the request and the buffers it keeps alive
        self._waiting = set()
        self._lock = threading.Lock()

    def submit(self, n, limbs, q, a, b):
        req, c, keep = _request(n, limbs, q, a, b)
        ticket = ctypes.c_uint64()
        _check(self.lib, self.lib.pm_server_submit(self.handle, ctypes.byref(req), ctypes.byref(ticket)))
        self._pending[ticket.value] = (req, c, keep)
        return ticket.value

    def wait(self, ticket):
        # A ticket is waited once; its buffers stay referenced until that wait
        # has returned, so a second waiter must not release them.
        with self._lock:
            if ticket not in self._pending or ticket in self._waiting:
                raise PolymulError(PM_ERR_INVALID_ARGUMENT, "unknown ticket or one already being waited for")
            self._waiting.add(ticket)
        times = RequestTimes()
        try:
            status = self.lib.pm_server_wait(self.handle, ticket, ctypes.byref(times))
        finally:
            with self._lock:
                self._waiting.discard(ticket)
        with self._lock:
            _, c, _ = self._pending.pop(ticket)
        _check(self.lib, status)
        return c, {"queue_ms": times.queue_ms, "service_ms": times.service_ms, "total_ms": times.total_ms}

    def close(self):
        if getattr(self, "handle", None):
            self.lib.pm_server_destroy(self.handle)
            self.handle = None
            self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
