"""Trace output in the reference's RayTracingExodus layout: the writer's file
has the gold file's dimensions, variables, attributes and element-variable
names (tests/golden/gold_rays_layout.json, taken from the reference's
test/tests/study/gold/openmc_study_rays.e), nodes shared between consecutive
visits, and the values it was given; its seed helper reproduces the gold's
source seeds."""
import json
import os

import numpy as np
from scipy.io import netcdf_file

from macaw_amd import abi
from macaw_amd.ray_output import ELEM_VARS, init_seed, write_rays_exodus

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_source_seeds_match_gold():
    d = json.load(open(os.path.join(GOLD, "gold_rays.json")))
    firsts = [r for r in d["records"] if r["intersections"] == 0]
    assert len(firsts) == 30
    for r in firsts:
        # the reference stores seeds as doubles
        assert int(float(init_seed(r["ray_id"] + 1, 1))) == r["seed_source"], r["ray_id"]


def synthetic_trace():
    rng = np.random.default_rng(3)
    visits = [4, 1, 2]
    t = np.zeros(sum(visits), abi.trace_dtype())
    k = 0
    for h, nv in enumerate(visits):
        p = rng.uniform(-1, 1, 3)
        for v in range(nv):
            q = p + rng.uniform(-0.1, 0.1, 3)
            t[k] = (h * 5, v, 100 + k, p, q, 1e6 - 10 * k, 2**63 + 7 * k, abi.EVENT_CROSS, 0)
            p = q
            k += 1
    t["event"][[3, 4, 6]] = [abi.EVENT_LEAK, abi.EVENT_ABSORB, abi.EVENT_LEAK]
    return t


def test_writer_layout_matches_gold(tmp_path):
    lay = json.load(open(os.path.join(GOLD, "gold_rays_layout.json")))
    t = synthetic_trace()
    path = str(tmp_path / "rays.e")
    counts = write_rays_exodus(path, t, batch=3, n_particles=30)
    assert counts == {"elements": 7, "nodes": 10, "rays": 3}
    f = netcdf_file(path, "r", mmap=False)
    for k, v in lay["dimensions_fixed"].items():
        assert f.dimensions[k] == v, k
    assert [k for k, v in f.dimensions.items() if v is None] == lay["unlimited"]
    assert set(f.variables) == set(lay["variables"])
    for k, (dims, tc) in lay["variables"].items():
        assert list(f.variables[k].dimensions) == dims and f.variables[k].typecode() == tc, k
    for k, attrs in lay["variable_attributes"].items():
        got = {a: (b.decode() if isinstance(b, bytes) else b) for a, b in f.variables[k]._attributes.items()}
        assert got == attrs, k
    names = [b"".join(x).decode().strip("\x00 ") for x in f.variables["name_elem_var"][:]]
    assert names == lay["elem_var_names"] == ELEM_VARS
    for k, v in lay["attributes"].items():
        assert np.isclose(float(f._attributes[k]), v), k
    assert f.dimensions["num_nodes"] == f.dimensions["num_elem"] + 3
    assert f.variables["time_whole"][0] == 3.0
    # consecutive visits of a ray share a node; rays do not
    con = f.variables["connect1"][:] - 1
    X = np.stack([f.variables[f"coord{a}"][:] for a in "xyz"], 1)
    assert np.array_equal(X[con[:, 0]], t["start"]) and np.array_equal(X[con[:, 1]], t["end"])
    assert list(con[:, 0]) == [0, 1, 2, 3, 5, 7, 8]
    vals = {n: f.variables[f"vals_elem_var{i + 1}eb1"][0] for i, n in enumerate(names)}
    assert np.array_equal(vals["ray_id"], 60 + t["history"]) and np.array_equal(vals["aux_id"], t["history"] + 1)
    assert np.array_equal(vals["intersections"], t["visit"]) and np.array_equal(vals["aux_energy"], t["energy"])
    assert np.array_equal(vals["aux_seed_tracking"], t["seed_tracking"].astype(np.float64))
    assert np.all(vals["aux_weight"] == 1.0) and np.all(vals["aux_particle_type"] == 0.0)
    assert vals["aux_seed_source"][0] == float(init_seed(61, 1))
    f.close()


def test_cpp_writer_equals_python_writer(tmp_path):
    """The driver's C++ writer and the Python writer produce the same file
    contents (variables, dimensions, attributes) for the same records."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "harness")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(root, "include"),
                        os.path.join(root, "tests", "exodus_harness.cpp"),
                        os.path.join(root, "macaw_amd", "host", "exodus.cpp"), "-o", exe],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    t = synthetic_trace()
    rec = str(tmp_path / "rec.bin")
    with open(rec, "wb") as fh:
        fh.write(np.int64(t.size).tobytes())
        fh.write(t.tobytes())
    a, b = str(tmp_path / "cpp.e"), str(tmp_path / "py.e")
    r = subprocess.run([exe, rec, a, "3", "30", "1"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    write_rays_exodus(b, t, batch=3, n_particles=30, master_seed=1)
    fa, fb = netcdf_file(a, "r", mmap=False), netcdf_file(b, "r", mmap=False)
    assert dict(fa.dimensions) == dict(fb.dimensions)
    assert set(fa.variables) == set(fb.variables)
    for k in fa.variables:
        va, vb = fa.variables[k], fb.variables[k]
        assert va.dimensions == vb.dimensions and va.typecode() == vb.typecode(), k
        assert np.array_equal(va[:], vb[:]), k
        assert va._attributes == vb._attributes, k
    for k, v in fb._attributes.items():
        if k != "title":
            assert fa._attributes[k] == v, k
    fa.close()
    fb.close()
