"""Traced histories in the layout of MOOSE's RayTracingExodus output.

MaCaw's regression study writes its rays with

    [Outputs/rays]  type = RayTracingExodus  output_data = true  output_aux_data = true

(test/tests/study/openmc_study.i:73-81): an Exodus II file (NetCDF, 64-bit
offset) with one EDGE2 element per element visit of a ray, consecutive visits
of a ray sharing their node, and one element variable per Ray datum.  MaCaw
registers eight aux data (src/userobjects/OpenMCStudy.C:190-197); MOOSE adds
the visit index (`intersections`) and the ray id; the variables are named in
alphabetical order.  `write_rays_exodus` writes the visit records of
`mcw_trace_get` in that layout, so a trace can be read by the same tools as
the reference's gold file.

What the records carry, per visit: the element, the start and end points,
and the energy and tracking seed at the visit's end.  Weight (1), progeny (0)
and particle type (0, neutron) are the constants of an analog fixed-source
neutron run; the source and URR seeds are their initial values (no
unresolved-resonance treatment changes the URR stream here).
"""
from __future__ import annotations

import numpy as np

_M = 6364136223846793005
_C = 1442695040888963407
_MASK = (1 << 64) - 1
STRIDE = 152917

ELEM_VARS = ["aux_energy", "aux_id", "aux_n_progeny", "aux_particle_type", "aux_seed_URR",
             "aux_seed_source", "aux_seed_tracking", "aux_weight", "intersections", "ray_id"]


def future_seed(n: int, seed: int) -> int:
    """The LCG state n steps after `seed` (Brown 1994, arbitrary-stride skip
    ahead), as OpenMC's future_seed."""
    g, c, g_new, c_new = _M, _C, 1, 0
    n &= _MASK
    while n:
        if n & 1:
            g_new = (g_new * g) & _MASK
            c_new = (c_new * g + c) & _MASK
        c = ((g + 1) * c) & _MASK
        g = (g * g) & _MASK
        n >>= 1
    return (g_new * seed + c_new) & _MASK


def init_seed(particle_id: int, stream: int, master_seed: int = 1) -> int:
    """Initial seed of a particle's stream (0 tracking, 1 source, 2 URR)."""
    return future_seed(particle_id * STRIDE, master_seed + stream)


def _chars(rows, n):
    out = np.zeros((len(rows), n), "S1")
    for i, r in enumerate(rows):
        b = r.encode()[: n - 1]
        out[i, : len(b)] = np.frombuffer(b, "S1")
    return out


def write_rays_exodus(path, trace, batch: int, n_particles: int, master_seed: int = 1,
                      title: str = "rays") -> dict:
    """Write visit records (abi.trace_dtype(), ordered by history and visit)
    of batch `batch` with `n_particles` particles.  Ray ids are the particles'
    indices across batches, (batch - 1) n_particles + history, as MOOSE
    numbers them; aux_id is the 1-based index within the batch.  Returns the
    counts written."""
    from scipy.io import netcdf_file

    trace = np.asarray(trace)
    n_elem = int(trace.size)
    hist = trace["history"].astype(np.int64)
    starts = np.ones(n_elem, bool)
    if n_elem:
        starts[1:] = hist[1:] != hist[:-1]
    n_rays = int(starts.sum())
    n_nodes = n_elem + n_rays
    # nodes: every visit's start, plus the end of each ray's last visit
    ends = np.zeros(n_elem, bool)
    if n_elem:
        ends[:-1] = starts[1:]
        ends[-1] = True
    xyz = np.zeros((n_nodes, 3))
    node_of_start = np.arange(n_elem) + np.cumsum(starts) - 1
    xyz[node_of_start] = trace["start"]
    xyz[node_of_start[ends] + 1] = trace["end"][ends]
    connect = np.stack([node_of_start + 1, node_of_start + 2], 1).astype(np.int32)

    pid = (batch - 1) * n_particles + hist
    seed_src = np.array([float(init_seed(int(p) + 1, 1, master_seed)) for p in pid])
    seed_urr = np.array([float(init_seed(int(p) + 1, 2, master_seed)) for p in pid])
    vals = {
        "aux_energy": trace["energy"],
        "aux_id": (hist + 1).astype(np.float64),
        "aux_n_progeny": np.zeros(n_elem),
        "aux_particle_type": np.zeros(n_elem),
        "aux_seed_URR": seed_urr,
        "aux_seed_source": seed_src,
        "aux_seed_tracking": trace["seed_tracking"].astype(np.float64),
        "aux_weight": np.ones(n_elem),
        "intersections": trace["visit"].astype(np.float64),
        "ray_id": pid.astype(np.float64),
    }

    f = netcdf_file(path, "w", version=2)
    f.api_version = np.float32(5.22)
    f.version = np.float32(5.22)
    f.floating_point_word_size = np.int32(8)
    f.file_size = np.int32(1)
    f.int64_status = np.int32(0)
    f.title = title
    f.maximum_name_length = np.int32(32)
    # (the record dimension first: the writer needs it there; order carries no meaning)
    for name, size in (("time_step", None), ("len_string", 33), ("len_line", 81), ("four", 4),
                       ("len_name", 33), ("num_dim", 3), ("num_nodes", n_nodes), ("num_elem", n_elem),
                       ("num_el_blk", 1), ("num_el_in_blk1", n_elem), ("num_nod_per_el1", 2),
                       ("num_elem_var", len(ELEM_VARS))):
        f.createDimension(name, size)
    v = f.createVariable("time_whole", "d", ("time_step",))
    v[0] = float(batch)  # the Transient step that ran the batch
    f.createVariable("eb_status", "i", ("num_el_blk",))[:] = [1]
    v = f.createVariable("eb_prop1", "i", ("num_el_blk",))
    v[:] = [0]
    v.name = "ID"
    for k, axis in enumerate("xyz"):
        f.createVariable(f"coord{axis}", "d", ("num_nodes",))[:] = xyz[:, k]
    f.createVariable("eb_names", "c", ("num_el_blk", "len_name"))[:] = _chars([""], 33)
    f.createVariable("coor_names", "c", ("num_dim", "len_name"))[:] = _chars(["", "", ""], 33)
    f.createVariable("node_num_map", "i", ("num_nodes",))[:] = np.arange(1, n_nodes + 1, dtype=np.int32)
    v = f.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
    v[:] = connect
    v.elem_type = "EDGE2"
    f.createVariable("elem_num_map", "i", ("num_elem",))[:] = np.arange(1, n_elem + 1, dtype=np.int32)
    f.createVariable("name_elem_var", "c", ("num_elem_var", "len_name"))[:] = _chars(ELEM_VARS, 33)
    for k, name in enumerate(ELEM_VARS):
        v = f.createVariable(f"vals_elem_var{k + 1}eb1", "d", ("time_step", "num_el_in_blk1"))
        v[0, :] = vals[name]
    f.createVariable("elem_var_tab", "i", ("num_el_blk", "num_elem_var"))[:] = np.ones((1, len(ELEM_VARS)), np.int32)
    f.close()
    return {"elements": n_elem, "nodes": n_nodes, "rays": n_rays}
