"""The driver (MOOSE-shaped host objects over the C ABI) on the GPU: the
per-element values its OpenMCTallyAux objects write equal the oracle's
tallies for the same problem (1e-10 relative)."""
import csv
import os
import subprocess

import numpy as np
import pytest

from macaw_amd import abi, build, problems as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "inputs")


def run_driver(name, tmp_path, *overrides):
    base = str(tmp_path / "out")
    args = [build.DRIVER, "-i", os.path.join(INPUTS, name), "Outputs/file_base=" + base] + list(overrides)
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    d = str(tmp_path / "p.dump")
    r = subprocess.run(args + ["--dump-problem", d], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stderr
    return base, P.load_dump(d)


def read_col(path, col):
    with open(path) as f:
        return np.array([float(r[col]) for r in csv.DictReader(f)])


def close(a, b, rtol=1e-10):
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300))


def test_c1_driver_matches_oracle(tmp_path, oracle_lib):
    base, p = run_driver("c1_hex8_reflective.i", tmp_path, "MonteCarlo/particles=1500", "MonteCarlo/batches=2")
    p.tallies = P.flux_tallies(both=True)
    o = oracle_lib.Oracle(p)
    for b in (1, 2):
        o.run_batch(b, p.n_particles, threads=4)
    assert close(read_col(base + "_flux_tl.csv", "flux_tl"), o.tally(1))
    assert close(read_col(base + "_flux_coll.csv", "flux_coll"), o.tally(2))
    segs = read_col(base + ".csv", "segments")
    assert segs.size == 2 and np.all(segs > 0)


def test_tally_driver_matches_oracle(tmp_path, oracle_lib):
    base, p = run_driver("openmc_tally.i", tmp_path, "MonteCarlo/batches=2")
    edges = [1e-5, 1, 1e3, 2e7]
    p.tallies = [
        P.TallySpec(id=1, estimator=abi.EST_COLLISION,
                    filters=[{"type": abi.FILTER_UNIVERSE}, {"type": abi.FILTER_ENERGY, "edges": edges}],
                    scores=[abi.SCORE_FLUX, abi.SCORE_ABSORPTION]),
        P.TallySpec(id=2, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                    scores=[abi.SCORE_FLUX, abi.SCORE_TOTAL], nuclides=[0, 2, -1]),
    ]
    o = oracle_lib.Oracle(p)
    for b in (1, 2):
        o.run_batch(b, p.n_particles, threads=4)
    t1 = o.tally(1).reshape(2, 3, 2)        # universe, energy, score
    sub = p.mesh.subdomain
    want1 = np.array([t1[{0: 0, 1: 1}[s], :, 1].sum() for s in sub])
    assert close(read_col(base + "_tally.csv", "tally"), want1)
    t2 = o.tally(2).reshape(-1, 3, 2)       # cell, nuclide bin, score
    want2 = t2[:, 0, 1]                     # nuclide H1, score total
    assert close(read_col(base + "_cell_flux.csv", "cell_flux"), want2)
    # MaCaw's aux index score_bin * n_nuclides + nuc_bin (OpenMCTallyAux.C:116-119):
    # score total (1) of nuclide H1 (0) reads flat bin 1 * 3 + 0 of the
    # nuclide-major results, which holds nuclide bin 1 (Fe56), score total
    base2, _ = run_driver("openmc_tally.i", tmp_path, "MonteCarlo/batches=2",
                          "AuxKernels/cell_val/score_index_layout=macaw")
    want2m = o.tally(2).reshape(-1, 6)[:, 1 * 3 + 0]
    assert close(read_col(base2 + "_cell_flux.csv", "cell_flux"), want2m)
    assert not close(want2m, want2)


def test_driver_ray_output_matches_gold_layout_and_source_seeds(tmp_path):
    """[Outputs/rays] type = RayTracingExodus, 30 particles x 3 batches as the
    reference's study test: the file has the gold file's layout, rays
    60..89 of batch 3, and every ray's source seed equals the gold's for the
    same ray (source seeds depend only on batch, particle count and id)."""
    import json
    from scipy.io import netcdf_file
    base = str(tmp_path / "out")
    args = [build.DRIVER, "-i", os.path.join(INPUTS, "c1_hex8_reflective.i"), "Outputs/file_base=" + base,
            "MonteCarlo/particles=30", "MonteCarlo/batches=3", "Outputs/rays/type=RayTracingExodus",
            "Outputs/rays/study=study", "Outputs/rays/execute_on=final"]
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    gold_dir = os.path.join(ROOT, "tests", "golden")
    lay = json.load(open(os.path.join(gold_dir, "gold_rays_layout.json")))
    gold = json.load(open(os.path.join(gold_dir, "gold_rays.json")))["records"]
    src_gold = {g["ray_id"]: g["seed_source"] for g in gold if g["intersections"] == 0}
    f = netcdf_file(base + "_rays.e", "r", mmap=False)
    for k, v in lay["dimensions_fixed"].items():
        assert f.dimensions[k] == v
    assert set(f.variables) == set(lay["variables"])
    names = [b"".join(x).decode().strip("\x00 ") for x in f.variables["name_elem_var"][:]]
    assert names == lay["elem_var_names"]
    vals = {n: f.variables[f"vals_elem_var{i + 1}eb1"][0] for i, n in enumerate(names)}
    ray = vals["ray_id"].astype(int)
    assert sorted(set(ray)) == list(range(60, 90))
    assert f.dimensions["num_nodes"] == f.dimensions["num_elem"] + 30
    assert f.variables["time_whole"][0] == 3.0
    for rid in range(60, 90):
        sel = ray == rid
        assert list(vals["intersections"][sel]) == list(range(sel.sum()))
        assert np.all(vals["aux_seed_source"][sel].astype(np.uint64) == np.uint64(src_gold[rid]))
        assert np.all(vals["aux_id"][sel] == rid - 59)
    f.close()


def test_eigen_driver_matches_oracle(tmp_path, oracle_lib):
    """run_mode = eigenvalue with one inactive batch (the reference study's
    settings): per-batch keff and the fission-power aux values equal the
    oracle's."""
    base, p = run_driver("openmc_study_eigen.i", tmp_path, "MonteCarlo/particles=400")
    p.tallies = [P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                             scores=[abi.SCORE_FISSION])]
    o = oracle_lib.Oracle(p, eigenvalue=True, n_inactive=p.n_inactive)
    kg, kf = [], []
    for b in (1, 2, 3):
        o.run_batch(b, p.n_particles, threads=4)
        k = o.keff()
        kg.append(k["k_generation"])
        kf.append(k["keff"])
    assert np.array_equal(read_col(base + ".csv", "k_generation"), np.array(kg))
    assert np.array_equal(read_col(base + ".csv", "keff"), np.array(kf))
    power = read_col(base + "_power.csv", "power")
    assert power.sum() > 0 and close(power, o.tally(1))


def test_temperature_driver_matches_oracle(tmp_path, oracle_lib):
    """Element temperatures from the temperature variable: per-batch keff and
    the fission-power aux values equal the oracle's."""
    base, p = run_driver("openmc_study_temperature.i", tmp_path, "MonteCarlo/particles=400")
    assert p.elem_temperature is not None and len(p.nuclide_temps) == 2
    p.tallies = [P.TallySpec(id=1, estimator=abi.EST_TRACKLENGTH, filters=[{"type": abi.FILTER_CELL}],
                             scores=[abi.SCORE_FISSION])]
    o = oracle_lib.Oracle(p, eigenvalue=True, n_inactive=p.n_inactive)
    kg = []
    for b in (1, 2, 3):
        o.run_batch(b, p.n_particles, threads=4)
        kg.append(o.keff()["k_generation"])
    assert np.array_equal(read_col(base + ".csv", "k_generation"), np.array(kg))
    power = read_col(base + "_power.csv", "power")
    assert power.sum() > 0 and close(power, o.tally(1))


def test_beavrs_assembly_driver_matches_oracle(tmp_path, oracle_lib):
    """The reference's BEAVRS assembly example (its Exodus mesh extruded to
    95,504 HEX8, a fission source on z = 100, reflective): per-batch keff and
    the kappa-fission power aux values equal the oracle's."""
    base, p = run_driver("beavrs_assembly_3d.i", tmp_path, "MonteCarlo/particles=20000", "MonteCarlo/batches=3")
    assert p.mesh.n_elem == 95504 and p.source.only_fissionable == 1
    p.tallies = [P.TallySpec(id=1, estimator=abi.EST_COLLISION, filters=[{"type": abi.FILTER_CELL}],
                             scores=[abi.SCORE_KAPPA_FISSION])]
    o = oracle_lib.Oracle(p, eigenvalue=True, n_inactive=p.n_inactive)
    kg = []
    for b in (1, 2, 3):
        o.run_batch(b, p.n_particles, threads=8)
        kg.append(o.keff()["k_generation"])
    assert np.array_equal(read_col(base + ".csv", "k_generation"), np.array(kg))
    power = read_col(base + "_power.csv", "power")
    assert power.sum() > 0 and close(power, o.tally(1))


def test_eigen_driver_ray_output_matches_gold_seeds(tmp_path):
    """The reference study's own settings -- eigenvalue, 30 particles, 3
    batches, 1 inactive -- with [Outputs/rays] type = RayTracingExodus: rays
    60..89 of batch 3, which start from resampled fission sites, carry the
    gold file's source seeds (they depend only on batch, particle count and
    id) and a visit sequence that starts at 0."""
    import json
    from scipy.io import netcdf_file
    base = str(tmp_path / "out")
    args = [build.DRIVER, "-i", os.path.join(INPUTS, "openmc_study_eigen.i"), "Outputs/file_base=" + base,
            "Outputs/rays/type=RayTracingExodus", "Outputs/rays/study=study", "Outputs/rays/execute_on=final"]
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    assert r.stdout.count("Keff ") == 3
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "gold_rays.json")))["records"]
    src_gold = {g["ray_id"]: g["seed_source"] for g in gold if g["intersections"] == 0}
    f = netcdf_file(base + "_rays.e", "r", mmap=False)
    names = [b"".join(x).decode().strip("\x00 ") for x in f.variables["name_elem_var"][:]]
    vals = {n: f.variables[f"vals_elem_var{i + 1}eb1"][0] for i, n in enumerate(names)}
    ray = vals["ray_id"].astype(int)
    assert sorted(set(ray)) == list(range(60, 90))
    for rid in range(60, 90):
        sel = ray == rid
        assert list(vals["intersections"][sel]) == list(range(sel.sum()))
        assert np.all(vals["aux_seed_source"][sel].astype(np.uint64) == np.uint64(src_gold[rid]))
    f.close()
