"""ctypes binding of libmacaw_amd.so (the C ABI in include/macaw_amd.h).

The library is the product: this module only loads it and declares the
signatures.  There is no fallback -- a missing or unloadable library is an
error.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .problems import Problem, load_into, nuclide_struct

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmacaw_amd.so")
_lib = None

# every symbol include/macaw_amd.h declares
EXPORTS = (
    "mcw_create", "mcw_destroy", "mcw_options_set", "mcw_last_error", "mcw_default_options", "mcw_mesh_set", "mcw_bc_set",
    "mcw_xs_set", "mcw_materials_set", "mcw_block_material_map", "mcw_source_set", "mcw_tally_add",
    "mcw_tally_shape", "mcw_run_batch", "mcw_tally_get", "mcw_tally_reset", "mcw_history_get",
    "mcw_server_create", "mcw_server_destroy", "mcw_server_last_error", "mcw_submit", "mcw_poll",
    "mcw_result", "mcw_trace_get", "mcw_debug_math_probe", "mcw_debug_xs_probe",
    "mcw_keff_get", "mcw_fission_bank_get", "mcw_source_bank_get", "mcw_xs_add_temperature",
    "mcw_temperature_set", "mcw_run_batch_noalloc", "mcw_server_tune", "mcw_debug_atomic_rate",
    "mcw_debug_gather_rate",
)


class McwError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise McwError(f"{LIB_PATH} is not built: run `python -m macaw_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u64, dp = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.POINTER(C.c_double)
    i32p, i64p, u8p, u64p = C.POINTER(i32), C.POINTER(i64), C.POINTER(C.c_uint8), C.POINTER(u64)
    sig = {
        "mcw_create": ([C.POINTER(vp), C.POINTER(abi.Options)], C.c_int),
        "mcw_destroy": ([vp], C.c_int),
        "mcw_options_set": ([vp, C.POINTER(abi.Options)], C.c_int),
        "mcw_last_error": ([vp], C.c_char_p),
        "mcw_default_options": ([C.POINTER(abi.Options)], None),
        "mcw_mesh_set": ([vp, i64, dp, i64, u8p, i64p, i32p, i32p, i32p], C.c_int),
        "mcw_bc_set": ([vp, i32, i32], C.c_int),
        "mcw_xs_set": ([vp, i32, C.POINTER(abi.Nuclide)], C.c_int),
        "mcw_xs_add_temperature": ([vp, i32, C.POINTER(abi.Nuclide)], C.c_int),
        "mcw_temperature_set": ([vp, i64, dp, C.c_double], C.c_int),
        "mcw_materials_set": ([vp, i32, C.POINTER(abi.Material)], C.c_int),
        "mcw_block_material_map": ([vp, i32, i32p, i32p], C.c_int),
        "mcw_source_set": ([vp, C.POINTER(abi.Source)], C.c_int),
        "mcw_tally_add": ([vp, i32, i32, i32, C.POINTER(abi.Filter), i32, i32p, i32, i32p], C.c_int),
        "mcw_tally_shape": ([vp, i32, i64p, i64p], C.c_int),
        "mcw_run_batch": ([vp, i64, i64, C.POINTER(abi.BatchStats)], C.c_int),
        "mcw_run_batch_noalloc": ([vp, i64, i64, C.POINTER(abi.BatchStats)], C.c_int),
        "mcw_tally_get": ([vp, i32, i32, dp, C.c_size_t], C.c_int),
        "mcw_tally_reset": ([vp], C.c_int),
        "mcw_history_get": ([vp, vp, C.c_size_t], C.c_int),
        "mcw_server_create": ([C.POINTER(vp), C.POINTER(vp), i32, i32, i64], C.c_int),
        "mcw_server_destroy": ([vp], C.c_int),
        "mcw_server_last_error": ([vp], C.c_char_p),
        "mcw_server_tune": ([vp, i64, C.c_double], C.c_int),
        "mcw_submit": ([vp, C.POINTER(abi.Request)], C.c_int),
        "mcw_poll": ([vp, i32p], C.c_int),
        "mcw_result": ([vp, C.POINTER(abi.RequestStatus), dp, C.c_size_t], C.c_int),
        "mcw_trace_get": ([vp, C.c_void_p, C.c_size_t, C.POINTER(C.c_int64)], C.c_int),
        "mcw_debug_math_probe": ([vp, dp, i64, dp, dp, dp, dp, u64p, dp, u64p], C.c_int),
        "mcw_debug_xs_probe": ([vp, i32, dp, i64, i32p, dp], C.c_int),
        "mcw_debug_atomic_rate": ([vp, i64, dp, dp], C.c_int),
        "mcw_debug_gather_rate": ([vp, i64, dp], C.c_int),
        "mcw_keff_get": ([vp, C.POINTER(abi.KeffResult)], C.c_int),
        "mcw_fission_bank_get": ([vp, C.c_void_p, C.c_size_t, C.POINTER(C.c_int64)], C.c_int),
        "mcw_source_bank_get": ([vp, C.c_void_p, C.c_size_t, C.POINTER(C.c_int64)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def default_options(**kw) -> abi.Options:
    o = abi.Options()
    lib().mcw_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Context:
    """One device context (mcw_ctx): a problem resident in HBM and its tallies."""

    def __init__(self, problem: Problem = None, **options):
        L = lib()
        self.L = L
        self.opts = default_options(**options)
        h = C.c_void_p()
        rc = L.mcw_create(C.byref(h), C.byref(self.opts))
        self.h = h
        if rc != 0:
            msg = L.mcw_last_error(h).decode() if h.value else "mcw_create failed"
            L.mcw_destroy(h)
            self.h = None
            raise McwError(msg)
        self.problem = None
        if problem is not None:
            self.load(problem)

    def _chk(self, rc):
        if rc != 0:
            raise McwError(self.L.mcw_last_error(self.h).decode())
        return rc

    def load(self, p: Problem):
        L, h = self.L, self.h
        calls = {
            "mesh_set": lambda *a: self._chk(L.mcw_mesh_set(h, *a)),
            "bc_set": lambda *a: self._chk(L.mcw_bc_set(h, *a)),
            "xs_set": lambda *a: self._chk(L.mcw_xs_set(h, *a)),
            "xs_add_temperature": lambda *a: self._chk(L.mcw_xs_add_temperature(h, *a)),
            "temperature_set": lambda *a: self._chk(L.mcw_temperature_set(h, *a)),
            "materials_set": lambda *a: self._chk(L.mcw_materials_set(h, *a)),
            "block_material_map": lambda *a: self._chk(L.mcw_block_material_map(h, *a)),
            "source_set": lambda *a: self._chk(L.mcw_source_set(h, *a)),
            "tally_add": lambda *a: self._chk(L.mcw_tally_add(h, *a)),
        }
        self._marshal = load_into(calls, p)
        self.problem = p
        return self

    def set_options(self, **kw):
        """Change schedule/tuning options in place (mcw_options_set)."""
        for k, v in kw.items():
            setattr(self.opts, k, v)
        self._chk(self.L.mcw_options_set(self.h, C.byref(self.opts)))

    def set_temperatures(self, kelvin, tolerance=10.0):
        """Element temperatures [K] in mesh order (None: every element reads
        each nuclide's first table), mcw_temperature_set."""
        if kelvin is None:
            self._chk(self.L.mcw_temperature_set(self.h, 0, None, float(tolerance)))
            return
        T = np.ascontiguousarray(kelvin, np.float64)
        self._chk(self.L.mcw_temperature_set(self.h, T.size, T.ctypes.data_as(C.POINTER(C.c_double)),
                                             float(tolerance)))

    def add_temperature(self, nuclide, table):
        """A further table (a NuclideData at table.temperature) of a nuclide."""
        keep = []
        n = nuclide_struct(table, keep)
        self._chk(self.L.mcw_xs_add_temperature(self.h, int(nuclide), C.byref(n)))

    def run_batch(self, batch, n_particles):
        st = abi.BatchStats()
        self._chk(self.L.mcw_run_batch(self.h, batch, n_particles, C.byref(st)))
        return st.as_dict()

    def run_batch_noalloc(self, batch, n_particles):
        """mcw_run_batch_noalloc: the same batch with no device allocation
        (the context must have run mcw_run_batch since it last changed)."""
        st = abi.BatchStats()
        self._chk(self.L.mcw_run_batch_noalloc(self.h, batch, n_particles, C.byref(st)))
        return st.as_dict()

    def tally_shape(self, tally_id):
        nf, ns = C.c_int64(), C.c_int64()
        self._chk(self.L.mcw_tally_shape(self.h, tally_id, C.byref(nf), C.byref(ns)))
        return nf.value, ns.value

    def tally(self, tally_id, kind=abi.RESULT_SUM):
        nf, ns = self.tally_shape(tally_id)
        out = np.zeros(nf * ns)
        self._chk(self.L.mcw_tally_get(self.h, tally_id, kind, out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def trace(self):
        """Visit records of the last batch (mcw_trace_get), ordered by
        (history, visit), as a numpy array of abi.trace_dtype()."""
        n = C.c_int64()
        self._chk(self.L.mcw_trace_get(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, abi.trace_dtype())
        if n.value:
            self._chk(self.L.mcw_trace_get(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def tally_reset(self):
        self._chk(self.L.mcw_tally_reset(self.h))

    def keff(self):
        """keff state after the last eigenvalue batch (mcw_keff_get) as a dict."""
        k = abi.KeffResult()
        self._chk(self.L.mcw_keff_get(self.h, C.byref(k)))
        return k.as_dict()

    def _sites(self, fn):
        n = C.c_int64()
        self._chk(fn(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, abi.site_dtype())
        if n.value:
            self._chk(fn(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def fission_bank(self):
        """The last eigenvalue batch's fission bank in (parent, progeny) order."""
        return self._sites(self.L.mcw_fission_bank_get)

    def source_bank(self):
        """The next eigenvalue batch's resampled source."""
        return self._sites(self.L.mcw_source_bank_get)

    def histories(self, n):
        out = np.zeros(n, abi.history_dtype())
        self._chk(self.L.mcw_history_get(self.h, out.ctypes.data, n))
        return out

    def math_probe(self, x, seeds):
        x = np.ascontiguousarray(x, np.float64)
        seeds = np.ascontiguousarray(seeds, np.uint64)
        n = x.size
        outs = [np.zeros(n) for _ in range(5)]
        so = np.zeros(n, np.uint64)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._chk(self.L.mcw_debug_math_probe(self.h, dp(x), n, dp(outs[0]), dp(outs[1]), dp(outs[2]), dp(outs[3]),
                                              seeds.ctypes.data_as(C.POINTER(C.c_uint64)), dp(outs[4]),
                                              so.ctypes.data_as(C.POINTER(C.c_uint64))))
        return {"log": outs[0], "sin": outs[1], "cos": outs[2], "sqrt": outs[3], "prn": outs[4], "future": so}

    def xs_probe(self, nuc, E):
        E = np.ascontiguousarray(E, np.float64)
        k = np.zeros(E.size, np.int32)
        xs = np.zeros((E.size, 4))
        self._chk(self.L.mcw_debug_xs_probe(self.h, nuc, E.ctypes.data_as(C.POINTER(C.c_double)), E.size,
                                            k.ctypes.data_as(C.POINTER(C.c_int32)),
                                            xs.ctypes.data_as(C.POINTER(C.c_double))))
        return k, xs

    def atomic_rate(self, n_bins=1 << 20):
        """(FP64 atomic adds per second into n_bins doubles, sum of the array
        after the probe = adds made) -- mcw_debug_atomic_rate."""
        r, t = C.c_double(), C.c_double()
        self._chk(self.L.mcw_debug_atomic_rate(self.h, int(n_bins), C.byref(r), C.byref(t)))
        return r.value, t.value

    def gather_rate(self, n_lines=1 << 22):
        """Reads of pseudo-random 128-byte lines per second from an array of
        n_lines lines -- mcw_debug_gather_rate."""
        r = C.c_double()
        self._chk(self.L.mcw_debug_gather_rate(self.h, int(n_lines), C.byref(r)))
        return r.value

    def close(self):
        if self.h is not None and self.h.value:
            self.L.mcw_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server over resident mesh contexts (mcw_server_*): continuous
    batching of fixed-source requests, each with a per-element track-length
    flux tally (per source particle)."""

    def __init__(self, contexts, max_inflight=8, bank_bytes=0):
        L = lib()
        self.L = L
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
        h = C.c_void_p()
        rc = L.mcw_server_create(C.byref(h), arr, len(self.contexts), max_inflight, bank_bytes)
        self.h = h
        if rc != 0:
            msg = L.mcw_server_last_error(h).decode() if h.value else "mcw_server_create failed"
            raise McwError(msg)
        self.n_elem = [c.problem.mesh.n_elem for c in self.contexts]
        self.mesh_of = {}

    def _chk(self, rc):
        if rc != 0:
            raise McwError(self.L.mcw_server_last_error(self.h).decode())

    def tune(self, round_histories=1 << 18, age_ms=250.0):
        """Histories per round and the age (ms) past which a request is served
        before shorter ones (mcw_server_tune; before the first submit)."""
        self._chk(self.L.mcw_server_tune(self.h, int(round_histories), float(age_ms)))

    def submit(self, request_id, mesh_slot, n_particles, source):
        r = abi.Request()
        r.request_id = request_id
        r.mesh_slot = mesh_slot
        r.n_particles = n_particles
        r.source = source
        self.mesh_of[request_id] = mesh_slot
        self._chk(self.L.mcw_submit(self.h, C.byref(r)))

    def poll(self):
        n = C.c_int32()
        self._chk(self.L.mcw_poll(self.h, C.byref(n)))
        return n.value

    def result(self, copy=True):
        """Pop one finished request: (status dict, flux in caller element
        order or None when copy is False)."""
        st = abi.RequestStatus()
        if copy:
            if getattr(self, "_buf", None) is None:
                self._buf = np.empty(max(self.n_elem))
            ptr = self._buf.ctypes.data_as(C.POINTER(C.c_double))
            self._chk(self.L.mcw_result(self.h, C.byref(st), ptr, self._buf.size))
        else:
            self._chk(self.L.mcw_result(self.h, C.byref(st), None, 0))
        d = {k: getattr(st, k) for k, _ in st._fields_}
        if not copy:
            return d, None
        return d, self._buf[:self.n_elem[self.mesh_of[st.request_id]]].copy()

    def run_all(self, sleep=0.0):
        """Poll until every submitted request has finished; returns {id: (status, flux)}."""
        import time
        out = {}
        want = len(self.mesh_of)
        while len(out) < want:
            n = self.poll()
            for _ in range(n):
                st, fl = self.result()
                out[st["request_id"]] = (st, fl)
            if sleep:
                time.sleep(sleep)
        return out

    def close(self):
        if self.h is not None and self.h.value:
            self.L.mcw_server_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
