"""ctypes mirror of the plain-data structs and enums in include/macaw_amd.h.

These are the types of the C ABI; the library loader (`macaw_amd.lib`) and
the test-side oracle loader share them so both are fed the same bytes.
"""
import ctypes as C

TET4, HEX8 = 4, 8
BC_VACUUM, BC_REFLECT = 0, 1
EST_ANALOG, EST_TRACKLENGTH, EST_COLLISION = 0, 1, 2
RESULT_VALUE, RESULT_SUM, RESULT_SUM_SQ = 0, 1, 2
FILTER_UNIVERSE, FILTER_CELL, FILTER_ENERGY, FILTER_PARTICLE = 1, 2, 3, 4
SCORE_FLUX, SCORE_TOTAL, SCORE_ABSORPTION, SCORE_SCATTER, SCORE_FISSION, SCORE_KAPPA_FISSION = 0, 1, 2, 3, 4, 5
SPACE_POINT, SPACE_BOX, SPACE_MESH_VOLUME = 0, 1, 2
ANGLE_ISOTROPIC, ANGLE_LINEAR_MU = 0, 1
ENERGY_MONO, ENERGY_WATT = 0, 1
MODE_PLAIN, MODE_BATCHED = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
FATE_ALIVE, FATE_LEAKED, FATE_ABSORBED, FATE_CUTOFF, FATE_LOST = 0, 1, 2, 3, 4
RUN_FIXED_SOURCE, RUN_EIGENVALUE = 0, 1
MATERIAL_BY_ID, MATERIAL_BY_INDEX = 0, 1
K_BOLTZMANN = 8.617333262e-5  # eV/K, as OpenMC

SCORE_NAMES = {"flux": SCORE_FLUX, "total": SCORE_TOTAL, "absorption": SCORE_ABSORPTION,
               "scatter": SCORE_SCATTER, "fission": SCORE_FISSION, "kappa-fission": SCORE_KAPPA_FISSION}
FILTER_NAMES = {"universe": FILTER_UNIVERSE, "cell": FILTER_CELL, "energy": FILTER_ENERGY,
                "particle": FILTER_PARTICLE}
ESTIMATOR_NAMES = {"ANALOG": EST_ANALOG, "TRACKLENGTH": EST_TRACKLENGTH,
                   "COLLISION": EST_COLLISION}

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("mode", C.c_int32), ("precision", C.c_int32),
                ("seed_roundtrip", C.c_int32), ("record_histories", C.c_int32),
                ("block_size", C.c_int32), ("bank_capacity", C.c_int64),
                ("max_segments", C.c_int64), ("master_seed", C.c_uint64),
                ("sort_cadence", C.c_int32), ("lanes_min", C.c_int32),
                ("waves_per_eu", C.c_int32), ("use_graph", C.c_int32), ("trace_histories", C.c_int32),
                ("run_mode", C.c_int32), ("n_inactive", C.c_int32), ("material_index", C.c_int32),
                ("tally_stage", C.c_int32)]


class Nuclide(C.Structure):
    _fields_ = [("n_points", C.c_int32), ("awr", C.c_double), ("energy", _dp), ("total", _dp),
                ("absorption", _dp), ("scatter", _dp), ("fission", _dp), ("nu", C.c_double * 2),
                ("watt", C.c_double * 2), ("temperature", C.c_double), ("kappa", C.c_double)]


class Material(C.Structure):
    _fields_ = [("id", C.c_int32), ("n_nuclides", C.c_int32), ("nuclide", _i32p),
                ("density", _dp)]


class Source(C.Structure):
    _fields_ = [("space", C.c_int32), ("angle", C.c_int32), ("energy", C.c_int32),
                ("only_fissionable", C.c_int32), ("point", C.c_double * 3), ("box_lo", C.c_double * 3),
                ("box_hi", C.c_double * 3), ("axis", C.c_double * 3), ("beta", C.c_double),
                ("E0", C.c_double), ("watt_a", C.c_double), ("watt_b", C.c_double)]


class Filter(C.Structure):
    _fields_ = [("type", C.c_int32), ("n_bins", C.c_int32), ("ids", _i64p), ("edges", _dp)]


class BatchStats(C.Structure):
    _fields_ = [("histories", C.c_int64), ("segments", C.c_int64), ("collisions", C.c_int64),
                ("absorptions", C.c_int64), ("leaks", C.c_int64), ("cutoffs", C.c_int64),
                ("lost", C.c_int64), ("reflections", C.c_int64), ("track_length", C.c_double),
                ("time_ms", C.c_double), ("kernel_ms", C.c_double),
                ("kernel_launches", C.c_int64), ("fp64_rechecks", C.c_int64), ("hex_segments", C.c_int64),
                ("trace_records", C.c_int64), ("stage_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TraceRecord(C.Structure):
    _fields_ = [("history", C.c_int64), ("visit", C.c_int32), ("element", C.c_int32),
                ("start", C.c_double * 3), ("end", C.c_double * 3), ("energy", C.c_double),
                ("seed_tracking", C.c_uint64), ("event", C.c_int32), ("pad", C.c_int32)]


def trace_dtype():
    """numpy dtype matching mcw_trace_record."""
    import numpy as np
    return np.dtype([("history", np.int64), ("visit", np.int32), ("element", np.int32),
                     ("start", np.float64, 3), ("end", np.float64, 3), ("energy", np.float64),
                     ("seed_tracking", np.uint64), ("event", np.int32), ("pad", np.int32)])


EVENT_CROSS, EVENT_REFLECT, EVENT_LEAK, EVENT_ABSORB, EVENT_CUTOFF, EVENT_LOST = 0, 1, 2, 4, 5, 6


class Site(C.Structure):
    _fields_ = [("r", C.c_double * 3), ("u", C.c_double * 3), ("E", C.c_double),
                ("parent", C.c_int64), ("progeny", C.c_int32), ("element", C.c_int32)]


def site_dtype():
    """numpy dtype matching mcw_site."""
    import numpy as np
    return np.dtype([("r", np.float64, 3), ("u", np.float64, 3), ("E", np.float64),
                     ("parent", np.int64), ("progeny", np.int32), ("element", np.int32)])


class KeffResult(C.Structure):
    _fields_ = [("batch", C.c_int64), ("fission_sites", C.c_int64), ("sites_dropped", C.c_int64),
                ("n_active", C.c_int32), ("pad", C.c_int32), ("k_generation", C.c_double),
                ("k_collision", C.c_double), ("k_absorption", C.c_double), ("keff", C.c_double),
                ("keff_std", C.c_double), ("keff_used", C.c_double), ("close_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class History(C.Structure):
    _fields_ = [("id", C.c_int64), ("digest", C.c_uint64), ("seed", C.c_uint64),
                ("E", C.c_double), ("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("n_segments", C.c_int32), ("n_collisions", C.c_int32), ("elem", C.c_int32),
                ("fate", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("request_id", C.c_int64), ("mesh_slot", C.c_int32), ("pad", C.c_int32),
                ("n_particles", C.c_int64), ("source", Source)]


class RequestStatus(C.Structure):
    _fields_ = [("request_id", C.c_int64), ("histories", C.c_int64), ("segments", C.c_int64),
                ("collisions", C.c_int64), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double)]


HISTORY_DTYPE = None


def history_dtype():
    """numpy dtype matching mcw_history (for bulk copies)."""
    import numpy as np
    global HISTORY_DTYPE
    if HISTORY_DTYPE is None:
        HISTORY_DTYPE = np.dtype([("id", "<i8"), ("digest", "<u8"), ("seed", "<u8"), ("E", "<f8"),
                                  ("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                                  ("n_segments", "<i4"), ("n_collisions", "<i4"),
                                  ("elem", "<i4"), ("fate", "<i4")])
        assert HISTORY_DTYPE.itemsize == C.sizeof(History)
    return HISTORY_DTYPE


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
