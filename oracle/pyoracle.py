"""ctypes loader for the CPU oracle (test infrastructure only).

Mirrors the C ABI of the product library (mcw_* -> orc_*) so parity tests
feed both the same Problem and compare tallies and per-history records.
"""
import ctypes as C
import os

import numpy as np

from macaw_amd import abi
from macaw_amd.problems import Problem, load_into

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64, dp = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(abi.Options)]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [vp]
        for name in ("orc_mesh_set", "orc_bc_set", "orc_xs_set", "orc_materials_set",
                     "orc_block_material_map", "orc_source_set", "orc_tally_add", "orc_tally_get",
                     "orc_run_batch", "orc_history_get", "orc_sample_sources", "orc_micro_xs"):
            getattr(L, name).restype = C.c_int
        u8p, i32p, i64p = C.POINTER(C.c_uint8), C.POINTER(i32), C.POINTER(i64)
        L.orc_mesh_set.argtypes = [vp, i64, dp, i64, u8p, i64p, i32p, i32p, i32p]
        L.orc_bc_set.argtypes = [vp, i32, i32]
        L.orc_xs_set.argtypes = [vp, i32, C.POINTER(abi.Nuclide)]
        L.orc_xs_add_temperature.restype = C.c_int
        L.orc_xs_add_temperature.argtypes = [vp, i32, C.POINTER(abi.Nuclide)]
        L.orc_temperature_set.restype = C.c_int
        L.orc_temperature_set.argtypes = [vp, i64, dp, C.c_double]
        L.orc_materials_set.argtypes = [vp, i32, C.POINTER(abi.Material)]
        L.orc_block_material_map.argtypes = [vp, i32, i32p, i32p]
        L.orc_source_set.argtypes = [vp, C.POINTER(abi.Source)]
        L.orc_tally_add.argtypes = [vp, i32, i32, i32, C.POINTER(abi.Filter), i32, i32p, i32, i32p]
        L.orc_run_batch.argtypes = [vp, i64, i64, i32, i32, C.POINTER(abi.BatchStats)]
        L.orc_tally_get.argtypes = [vp, i32, i32, dp, C.c_size_t]
        L.orc_history_get.argtypes = [vp, C.c_void_p, C.c_size_t]
        L.orc_trace_history.restype = i64
        L.orc_trace_history.argtypes = [vp, i64, i64, i64, C.POINTER(i32), C.POINTER(i32), dp, i64]
        L.orc_mesh_n_faces.restype = i64
        L.orc_mesh_n_faces.argtypes = [vp]
        L.orc_mesh_planes.argtypes = [vp, dp]
        L.orc_mesh_neighbors.argtypes = [vp, C.POINTER(i64), C.POINTER(i32)]
        L.orc_mesh_n_splits.argtypes = [vp]
        L.orc_mesh_n_splits.restype = i64
        L.orc_mesh_splits.argtypes = [vp, C.POINTER(i64), C.POINTER(C.c_double), C.POINTER(i64), C.POINTER(i32)]
        L.orc_mesh_cdf.argtypes = [vp, dp]
        L.orc_locate.restype = i64
        L.orc_locate.argtypes = [vp, dp, C.c_double]
        L.orc_sample_sources.argtypes = [vp, i64, i64, i64, dp]
        L.orc_micro_xs.argtypes = [vp, i32, C.c_double, dp]
        L.orc_sample_nuclide_x.restype = i32
        L.orc_sample_nuclide_x.argtypes = [vp, i32, C.c_double, C.POINTER(u64)]
        L.orc_macro_total.restype = C.c_double
        L.orc_macro_total.argtypes = [vp, i32, C.c_double]
        L.orc_log_x.restype = C.c_double
        L.orc_log_x.argtypes = [C.c_double]
        L.orc_sincos_2pi_x.argtypes = [C.c_double, dp, dp]
        L.orc_prn_x.restype = C.c_double
        L.orc_prn_x.argtypes = [C.POINTER(u64)]
        L.orc_step_x.restype = u64
        L.orc_step_x.argtypes = [u64]
        L.orc_future_seed_x.restype = u64
        L.orc_future_seed_x.argtypes = [u64, u64]
        L.orc_init_seed_x.restype = u64
        L.orc_init_seed_x.argtypes = [i64, i32, u64]
        L.orc_seed_roundtrip_x.restype = u64
        L.orc_seed_roundtrip_x.argtypes = [u64]
        L.orc_rotate_angle_x.argtypes = [dp, C.c_double, C.c_double]
        L.orc_elastic_x.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp]
        sp = C.POINTER(abi.Site)
        for name in ("orc_keff_get", "orc_fission_bank_get", "orc_source_bank_get", "orc_synchronize_bank"):
            getattr(L, name).restype = C.c_int
        L.orc_keff_get.argtypes = [vp, C.POINTER(abi.KeffResult)]
        L.orc_fission_bank_get.argtypes = [vp, C.c_void_p, C.c_size_t, C.POINTER(i64)]
        L.orc_source_bank_get.argtypes = [vp, C.c_void_p, C.c_size_t, C.POINTER(i64)]
        L.orc_tree_sum.restype = C.c_double
        L.orc_tree_sum.argtypes = [dp, i64]
        L.orc_synchronize_bank.argtypes = [C.c_void_p, i64, i64, i64, u64, C.c_void_p]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def tree_sum(v):
    """The fixed-order sum of the per-history keff scores."""
    v = np.ascontiguousarray(v, np.float64)
    return lib().orc_tree_sum(v.ctypes.data_as(C.POINTER(C.c_double)), v.size)


def synchronize_bank(bank, batch, n, master_seed=1):
    """OpenMCStudy::synchronizeBanks of `bank` (abi.site_dtype()) to n sites."""
    bank = np.ascontiguousarray(bank, abi.site_dtype())
    out = np.zeros(n, abi.site_dtype())
    if lib().orc_synchronize_bank(bank.ctypes.data, bank.size, batch, n, master_seed, out.ctypes.data) != 0:
        raise OracleError("synchronize_bank: empty bank or n < 1")
    return out


class Oracle:
    """History-based CPU transport of a Problem."""

    def __init__(self, problem: Problem = None, seed_roundtrip=False, max_segments=0, master_seed=1,
                 eigenvalue=False, n_inactive=0, material_index=abi.MATERIAL_BY_ID):
        L = lib()
        o = abi.Options()
        o.material_index = material_index
        o.seed_roundtrip = int(seed_roundtrip)
        o.max_segments = max_segments
        o.master_seed = master_seed
        o.run_mode = abi.RUN_EIGENVALUE if eigenvalue else abi.RUN_FIXED_SOURCE
        o.n_inactive = n_inactive
        self.h = L.orc_create(C.byref(o))
        self.L = L
        self.problem = None
        if problem is not None:
            self.load(problem)

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(self.L.orc_last_error(self.h).decode())
        return rc

    def load(self, p: Problem):
        L, h = self.L, self.h
        calls = {
            "mesh_set": lambda *a: self._chk(L.orc_mesh_set(h, *a)),
            "bc_set": lambda *a: self._chk(L.orc_bc_set(h, *a)),
            "xs_set": lambda *a: self._chk(L.orc_xs_set(h, *a)),
            "xs_add_temperature": lambda *a: self._chk(L.orc_xs_add_temperature(h, *a)),
            "temperature_set": lambda *a: self._chk(L.orc_temperature_set(h, *a)),
            "materials_set": lambda *a: self._chk(L.orc_materials_set(h, *a)),
            "block_material_map": lambda *a: self._chk(L.orc_block_material_map(h, *a)),
            "source_set": lambda *a: self._chk(L.orc_source_set(h, *a)),
            "tally_add": lambda *a: self._chk(L.orc_tally_add(h, *a)),
        }
        self._marshal = load_into(calls, p)
        self.problem = p
        self.n_elem = p.mesh.n_elem
        return self

    def run_batch(self, batch, n_particles, threads=0, record=False):
        st = abi.BatchStats()
        self._chk(self.L.orc_run_batch(self.h, batch, n_particles, threads, int(record), C.byref(st)))
        return st.as_dict()

    def histories(self, n):
        out = np.zeros(n, abi.history_dtype())
        self._chk(self.L.orc_history_get(self.h, out.ctypes.data, n))
        return out

    def keff(self):
        k = abi.KeffResult()
        self._chk(self.L.orc_keff_get(self.h, C.byref(k)))
        return k.as_dict()

    def _sites(self, fn):
        n = C.c_int64(0)
        self._chk(fn(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, abi.site_dtype())
        self._chk(fn(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def fission_bank(self):
        return self._sites(self.L.orc_fission_bank_get)

    def source_bank(self):
        return self._sites(self.L.orc_source_bank_get)

    def tally(self, tally_id, kind=abi.RESULT_SUM, size=None):
        t = [t for t in self.problem.tallies if t.id == tally_id][0]
        n = size if size is not None else self._tally_size(t)
        out = np.zeros(n)
        self._chk(self.L.orc_tally_get(self.h, tally_id, kind, out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out

    def _tally_size(self, t):
        nf = 1
        for f in t.filters:
            if f["type"] == abi.FILTER_CELL:
                nf *= len(f["ids"]) if f.get("ids") is not None else self.n_elem
            elif f["type"] == abi.FILTER_UNIVERSE:
                nf *= len(f["ids"]) if f.get("ids") is not None else len(np.unique(self.problem.mesh.subdomain))
            elif f["type"] == abi.FILTER_ENERGY:
                nf *= len(f["edges"]) - 1
        return nf * len(t.scores) * len(t.nuclides)

    def trace(self, batch, n_particles, i, max_n=1 << 20):
        el = np.zeros(max_n, np.int32)
        ev = np.zeros(max_n, np.int32)
        le = np.zeros(max_n)
        n = self.L.orc_trace_history(self.h, batch, n_particles, i, el.ctypes.data_as(C.POINTER(C.c_int32)),
                                     ev.ctypes.data_as(C.POINTER(C.c_int32)),
                                     le.ctypes.data_as(C.POINTER(C.c_double)), max_n)
        if n < 0:
            raise OracleError(self.L.orc_last_error(self.h).decode())
        n = min(n, max_n)
        return el[:n], ev[:n], le[:n]

    def sources(self, batch, n_particles, n):
        out = np.zeros((n, 8))
        self._chk(self.L.orc_sample_sources(self.h, batch, n_particles, n,
                                            out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def planes(self):
        nf = self.L.orc_mesh_n_faces(self.h)
        out = np.zeros((nf, 4))
        self.L.orc_mesh_planes(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def neighbors(self):
        nf = self.L.orc_mesh_n_faces(self.h)
        nb = np.zeros(nf, np.int64)
        sd = np.zeros(nf, np.int32)
        self.L.orc_mesh_neighbors(self.h, nb.ctypes.data_as(C.POINTER(C.c_int64)),
                                  sd.ctypes.data_as(C.POINTER(C.c_int32)))
        return nb, sd

    def splits(self):
        """(split index per side slot, -1 = none; planes (k, 4); elements (k, 2); sides (k, 2))."""
        nf = self.L.orc_mesh_n_faces(self.h)
        k = self.L.orc_mesh_n_splits(self.h)
        so = np.zeros(nf, np.int64)
        pl = np.zeros((max(k, 1), 4))
        el = np.zeros((max(k, 1), 2), np.int64)
        sd = np.zeros((max(k, 1), 2), np.int32)
        self.L.orc_mesh_splits(self.h, so.ctypes.data_as(C.POINTER(C.c_int64)),
                               pl.ctypes.data_as(C.POINTER(C.c_double)), el.ctypes.data_as(C.POINTER(C.c_int64)),
                               sd.ctypes.data_as(C.POINTER(C.c_int32)))
        return so, pl[:k], el[:k], sd[:k]

    def cdf(self):
        out = np.zeros(self.n_elem)
        self.L.orc_mesh_cdf(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def sample_nuclides(self, mat, E, seeds):
        """The nuclide each seed picks in a collision in material `mat` at E."""
        out = np.empty(len(seeds), np.int32)
        for i, sd in enumerate(seeds):
            s = C.c_uint64(int(sd))
            out[i] = self.L.orc_sample_nuclide_x(self.h, mat, float(E), C.byref(s))
        return out

    def micro_xs(self, nuc, E):
        out = np.zeros(4)
        self.L.orc_micro_xs(self.h, nuc, E, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass


# --- elementary functions ------------------------------------------------------

def log(x):
    return lib().orc_log_x(x)


def sincos_2pi(xi):
    s, c = C.c_double(), C.c_double()
    lib().orc_sincos_2pi_x(xi, C.byref(s), C.byref(c))
    return s.value, c.value


def prn(seed):
    s = C.c_uint64(seed)
    v = lib().orc_prn_x(C.byref(s))
    return v, s.value


def step(s):
    return lib().orc_step_x(s)


def future_seed(n, s):
    return lib().orc_future_seed_x(n, s)


def init_seed(i, stream, master=1):
    return lib().orc_init_seed_x(i, stream, master)


def seed_roundtrip(s):
    return lib().orc_seed_roundtrip_x(s)


def rotate_angle(u, mu, xi):
    a = np.ascontiguousarray(u, np.float64).copy()
    lib().orc_rotate_angle_x(a.ctypes.data_as(C.POINTER(C.c_double)), mu, xi)
    return a


def elastic(E, A, mu_cm):
    ep, ml = C.c_double(), C.c_double()
    lib().orc_elastic_x(E, A, mu_cm, C.byref(ep), C.byref(ml))
    return ep.value, ml.value
