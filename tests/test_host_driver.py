"""The MOOSE-shaped host layer (macaw_amd/host) and its driver on the CPU:
inputs laid out as the reference's test inputs are read and validated, the
reference's parameter errors are reproduced, and the problem the driver hands
to the library can be replayed on the oracle."""
import os
import subprocess

import numpy as np
import pytest

from macaw_amd import abi, build, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "inputs")
DRIVER = build.DRIVER


@pytest.fixture(scope="module", autouse=True)
def driver_built():
    if not os.path.exists(DRIVER):
        build.build_driver()


def run(args, cwd=None):
    return subprocess.run([DRIVER] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120)


@pytest.mark.parametrize("name", ["c1_hex8_reflective.i", "openmc_tally.i", "openmc_study_eigen.i",
                                  "openmc_study_temperature.i", "beavrs_assembly_3d.i"])
def test_inputs_validate(name):
    r = run(["-i", os.path.join(INPUTS, name), "--check-input"])
    assert r.returncode == 0, r.stderr
    assert "input ok" in r.stdout


ERRORS = [
    ("c1_hex8_reflective.i", "UserObjects/tracklength/particle_type=photon", "Only neutrons are currently supported."),
    ("c1_hex8_reflective.i", "UserObjects/collision/estimator=ANALOG",
     "Only collision and tracklength estimators are currently supported"),
    ("c1_hex8_reflective.i", "UserObjects/collision/execute_on=timestep_end",
     "execute_on must be INITIAL to ensure tallies are created once at the beginning of the simulation"),
    ("c1_hex8_reflective.i", "UserObjects/collision/filters=energy",
     "The energy bins parameter should be supplied if an energy filter is requested."),
    ("c1_hex8_reflective.i", "RayKernels/collision/blocks=0 1",
     "The blocks parameter must be the same size as the materials parameter."),
    ("c1_hex8_reflective.i", "Variables/temperature/order=FIRST",
     "Only CONST MONOMIAL temperatures (current order: 1) are supported."),
    ("c1_hex8_reflective.i", "RayBCs/reflect/boundary=back nowhere", "the boundary 'nowhere' was not found"),
    ("c1_hex8_reflective.i", "MonteCarlo/Materials/graphite/nuclides=U235", "unknown nuclide 'U235'"),
    ("openmc_study_eigen.i", "MonteCarlo/run_mode=criticality", "run_mode must be eigenvalue or fixed source"),
    ("openmc_study_eigen.i", "MonteCarlo/inactive=3", "an eigenvalue run needs more batches than inactive batches"),
    ("openmc_study_eigen.i", "MonteCarlo/mode=plain", "eigenvalue runs need mode = batched"),
    ("openmc_study_eigen.i", "MonteCarlo/Nuclides/U235/fission_fraction=1.5", "the fission fraction must be in [0, 1]"),
    ("openmc_study_eigen.i", "MonteCarlo/Nuclides/U235/nu=1 2 3", "nu takes one or two coefficients"),
    ("openmc_study_eigen.i", "MonteCarlo/source/upper_right=0 0", "a box corner needs 3 coordinates"),
    ("openmc_study_temperature.i", "MonteCarlo/Nuclides/U235/temperatures=600 -1", "temperatures must be positive"),
    ("openmc_study_temperature.i", "MonteCarlo/Nuclides/O16/temperatures=600 600", "a temperature is given twice"),
    ("openmc_study_temperature.i", "ICs/centre_temperature/type=FunctionIC", "only ConstantIC sets the temperature"),
    ("openmc_study_temperature.i", "MonteCarlo/temperature_tolerance=-1", "the tolerance must be >= 0 K"),
]


@pytest.mark.parametrize("name,override,message", ERRORS)
def test_reference_parameter_errors(name, override, message):
    r = run(["-i", os.path.join(INPUTS, name), override, "--check-input"])
    assert r.returncode == 1
    assert message in r.stderr, r.stderr


RAY_OUTPUT = ["Outputs/rays/type=RayTracingExodus", "Outputs/rays/study=study", "Outputs/rays/execute_on=final"]


def test_ray_output_block_validates():
    r = run(["-i", os.path.join(INPUTS, "c1_hex8_reflective.i")] + RAY_OUTPUT + ["--check-input"])
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("override,message", [
    ("Outputs/rays/type=Exodus", "output type 'Exodus' is not supported"),
    ("Outputs/rays/execute_on=timestep_end", "only execute_on = final is supported for ray output"),
    ("Outputs/rays/study=other", "no study 'other'"),
])
def test_ray_output_block_errors(override, message):
    r = run(["-i", os.path.join(INPUTS, "c1_hex8_reflective.i")] + RAY_OUTPUT + [override, "--check-input"])
    assert r.returncode == 1
    assert message in r.stderr, r.stderr


def test_missing_material_is_reported_at_initial_setup(tmp_path):
    # CollisionKernel::initialSetup (CollisionKernel.C:84-90): a material id
    # that does not exist; found before the device is touched.
    r = run(["-i", os.path.join(INPUTS, "c1_hex8_reflective.i"), "RayKernels/collision/materials=7",
             "Outputs/file_base=" + str(tmp_path / "x")])
    assert r.returncode == 1
    assert "Could not find material 7 in OpenMC materials." in r.stderr


def test_hit_syntax(tmp_path):
    text = open(os.path.join(INPUTS, "c1_hex8_reflective.i")).read()
    text = text.replace("boundary = 'back front top right left bottom'",
                        "boundary = 'back front\n                top right\n left bottom'  # a comment")
    text = text.replace("[Executioner]\n  type = Transient\n[]", "[Executioner]\n  type = Transient\n[../]")
    f = tmp_path / "x.i"
    f.write_text(text)
    r = run(["-i", str(f), "--check-input"])
    assert r.returncode == 0, r.stderr
    bad = tmp_path / "bad.i"
    bad.write_text("[Mesh]\n  [gmg]\n    type = GeneratedMeshGenerator\n  []\n")
    r = run(["-i", str(bad), "--check-input"])
    assert r.returncode == 1 and "is not closed" in r.stderr


def test_eigen_dump_replays_on_the_oracle(tmp_path, oracle_lib):
    d = tmp_path / "eig.dump"
    r = run(["-i", os.path.join(INPUTS, "openmc_study_eigen.i"), "MonteCarlo/particles=200", "--dump-problem", str(d)])
    assert r.returncode == 0, r.stderr
    p = P.load_dump(str(d))
    assert p.run_mode == abi.RUN_EIGENVALUE and p.n_inactive == 1 and p.n_batches == 3
    u235 = p.nuclides[0]
    assert u235.fission is not None and u235.nu == (2.43, 0.0) and u235.watt == (0.988e6, 2.249e-6)
    assert np.allclose(u235.fission, 0.85 * u235.absorption, rtol=1e-15)
    assert all(n.fission is None for n in p.nuclides[1:])
    assert p.source.space == abi.SPACE_BOX and list(p.source.box_hi) == [0.62992] * 3
    o = oracle_lib.Oracle(p, eigenvalue=True, n_inactive=p.n_inactive)
    for b in (1, 2, 3):
        o.run_batch(b, 200, threads=2)
    k = o.keff()
    assert k["n_active"] == 2 and k["fission_sites"] > 0 and 0.5 < k["keff"] < 3.0


def test_material_index_option_reaches_the_library(tmp_path):
    """[MonteCarlo] material_index = index (MaCaw's id - 1, CollisionKernel.C:110)
    is handed to the library; the default is by id; other values are refused."""
    c1 = os.path.join(INPUTS, "c1_hex8_reflective.i")
    for value, want in ((None, abi.MATERIAL_BY_ID), ("index", abi.MATERIAL_BY_INDEX), ("id", abi.MATERIAL_BY_ID)):
        d = tmp_path / f"m_{value}.dump"
        extra = [f"MonteCarlo/material_index={value}"] if value else []
        r = run(["-i", c1, "MonteCarlo/particles=10", "--dump-problem", str(d)] + extra)
        assert r.returncode == 0, r.stderr
        assert P.load_dump(str(d)).material_index == want
    r = run(["-i", c1, "MonteCarlo/material_index=position", "--check-input"])
    assert r.returncode != 0 and "material_index must be id or index" in r.stderr


def test_dump_replays_on_the_oracle(tmp_path, oracle_lib):
    d = tmp_path / "c1.dump"
    r = run(["-i", os.path.join(INPUTS, "c1_hex8_reflective.i"), "MonteCarlo/particles=300", "--dump-problem", str(d)])
    assert r.returncode == 0, r.stderr
    p = P.load_dump(str(d))
    assert p.mesh.n_elem == 512 and p.n_particles == 300 and p.block_material == {0: 1}
    assert p.bcs == {b: abi.BC_REFLECT for b in range(6)}
    assert p.nuclides[0].energy.size == 20000 and p.nuclides[0].awr == 12.0
    p.tallies = P.flux_tallies(both=True)
    o = oracle_lib.Oracle(p)
    st = o.run_batch(1, 300, threads=2)
    assert st["lost"] == 0 and st["leaks"] == 0 and st["absorptions"] + st["cutoffs"] == 300
    assert o.tally(1).sum() > 0


def test_temperature_dump_replays_on_the_oracle(tmp_path, oracle_lib):
    """Element temperatures from the temperature variable (initial condition,
    then a block ConstantIC) and the nuclides' tables reach the library; on
    the oracle the run equals the same problem with each element's nearest
    tables made nuclides of their own."""
    from test_oracle_temperature import assert_same_histories, expand
    d = tmp_path / "temp.dump"
    r = run(["-i", os.path.join(INPUTS, "openmc_study_temperature.i"), "MonteCarlo/particles=300",
             "--dump-problem", str(d)])
    assert r.returncode == 0, r.stderr
    p = P.load_dump(str(d))
    centre = p.mesh.subdomain == 1
    assert centre.sum() == 1 and np.all(p.elem_temperature[centre] == 1150.0)
    assert np.all(p.elem_temperature[~centre] == 600.0) and p.temperature_tolerance == np.inf
    assert [n.temperature for n in p.nuclides] == [293.6] * 3
    assert {q: [t.temperature for t in ts] for q, ts in p.nuclide_temps.items()} == {
        0: [600.0, 900.0, 1200.0], 1: [600.0, 900.0, 1200.0]}
    assert p.nuclide_temps[0][2].fission is not None and p.nuclide_temps[0][2].nu == p.nuclides[0].nu
    q, n_nuc, _ = expand(p)
    assert n_nuc == 4  # U235 at 600 K (fuel only), O16 at 600 K and 1200 K, H1
    runs = []
    for prob in (p, q):
        o = oracle_lib.Oracle(prob, eigenvalue=True, n_inactive=p.n_inactive)
        for b in (1, 2):
            o.run_batch(b, 300, threads=1, record=True)
        runs.append((o.keff(), o.histories(300)))
    assert_same_histories(runs[0][1], runs[1][1])
    assert runs[0][0]["keff"] == runs[1][0]["keff"]


def test_exodus_mesh_is_read_and_extruded(tmp_path):
    """FileMeshGenerator + TransformGenerator + FancyExtruderGenerator on the
    reference example's QUAD4 assembly mesh, checked against the same file
    read by scipy's NetCDF reader: nodes layer by layer, each hex's nodes
    from its quad's, blocks, and side-set ids carried to the swept sides."""
    from scipy.io import netcdf_file
    d = tmp_path / "bv.dump"
    r = run(["-i", os.path.join(INPUTS, "beavrs_assembly_3d.i"), "MonteCarlo/particles=100", "--dump-problem", str(d)])
    assert r.returncode == 0, r.stderr
    p = P.load_dump(str(d))
    f = netcdf_file(os.path.join(INPUTS, "31enr_16_in.e"), "r", mmap=False)
    V = f.variables
    x, y = V["coordx"][:].copy(), V["coordy"][:].copy()
    n2 = x.size
    quads, blocks = [], []
    for k, bid in enumerate(V["eb_prop1"][:]):
        c = V[f"connect{k + 1}"][:].astype(np.int64) - 1
        quads.append(c)
        blocks.append(np.full(c.shape[0], bid))
    quads, blocks = np.concatenate(quads), np.concatenate(blocks)
    e2 = quads.shape[0]
    m = p.mesh
    assert m.n_nodes == 5 * n2 and m.n_elem == 4 * e2 and np.all(m.elem_type == abi.HEX8)
    for j in range(5):
        lay = m.xyz[j * n2:(j + 1) * n2]
        assert np.array_equal(lay[:, 0], x) and np.array_equal(lay[:, 1], y)
        assert np.all(lay[:, 2] == -100.0 + 100.0 * j)
    conn = m.conn.reshape(-1, 8).astype(np.int64)
    for j in range(4):
        h = conn[j * e2:(j + 1) * e2]
        assert np.array_equal(np.sort(h[:, :4], 1), np.sort(quads, 1) + j * n2)
        assert np.array_equal(h[:, 4:], h[:, :4] + n2)
    assert np.array_equal(m.subdomain, np.tile(blocks, 4))
    # every hex has positive volume: its bottom quad winds counterclockwise about +z
    b = m.xyz[conn[:, :4]]
    area2 = ((b[:, 0, 0] - b[:, 2, 0]) * (b[:, 1, 1] - b[:, 3, 1]) - (b[:, 1, 0] - b[:, 3, 0]) * (b[:, 0, 1] - b[:, 2, 1]))
    assert np.all(area2 > 0)
    # side-set ids: each 2D side set's sides, once per layer; bottom 8 and top 9 once per column
    fb = m.face_boundary.reshape(-1, 6)
    ids, counts = np.unique(fb, return_counts=True)
    got = dict(zip(ids.tolist(), counts.tolist()))
    for k, sid in enumerate(V["ss_prop1"][:]):
        assert got[int(sid)] == 4 * V[f"elem_ss{k + 1}"].shape[0], sid
    assert got[8] == e2 and got[9] == e2
    assert np.all(fb[:e2, 0] == 8) and np.all(fb[3 * e2:, 5] == 9) and np.all(fb[e2:, 0] == -1)


def test_beavrs_assembly_replays_on_the_oracle(tmp_path, oracle_lib):
    """The assembly example runs on the oracle: fission-source sites sit on
    z = 100 in fuel, nothing is lost in the extruded pin-cell mesh, and the
    kappa-fission score is kappa times the fission score (U235 is the only
    fissionable nuclide)."""
    d = tmp_path / "bv.dump"
    r = run(["-i", os.path.join(INPUTS, "beavrs_assembly_3d.i"), "MonteCarlo/particles=400", "--dump-problem", str(d)])
    assert r.returncode == 0, r.stderr
    p = P.load_dump(str(d))
    assert p.source.only_fissionable == 1 and p.run_mode == abi.RUN_EIGENVALUE
    kappa = [n.kappa for n in p.nuclides]
    assert kappa[0] == 193.4e6 and all(k == 0.0 for k in kappa[1:])
    p.tallies = [P.TallySpec(id=1, estimator=abi.EST_COLLISION, filters=[{"type": abi.FILTER_CELL}],
                             scores=[abi.SCORE_KAPPA_FISSION, abi.SCORE_FISSION])]
    o = oracle_lib.Oracle(p, eigenvalue=True, n_inactive=0)
    src = o.sources(1, 400, 400)
    elem = src[:, 7].astype(np.int64)
    assert np.all(src[:, 2] == 100.0) and np.all(p.mesh.subdomain[elem] == 10010)  # fuel 3.1%
    for b in (1, 2):
        st = o.run_batch(b, 400, threads=4)
        assert st["lost"] == 0 and st["leaks"] == 0 and st["absorptions"] + st["cutoffs"] == 400
    k = o.keff()
    assert 0.8 < k["k_generation"] < 2.0
    t = o.tally(1).reshape(-1, 2)
    assert t[:, 1].sum() > 0
    np.testing.assert_allclose(t[:, 0], 193.4e6 * t[:, 1], rtol=1e-13, atol=0)


# Exodus II face node lists (1-based node positions within the element), as
# the Exodus II manual numbers element sides
EXO_FACES = {
    "HEX8": [(1, 2, 6, 5), (2, 3, 7, 6), (3, 4, 8, 7), (1, 5, 8, 4), (1, 4, 3, 2), (5, 6, 7, 8)],
    "TETRA": [(1, 2, 4), (2, 3, 4), (1, 4, 3), (1, 3, 2)],
}
# libMesh side node lists (0-based)
LM_FACES = {
    "HEX8": [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)],
    "TETRA": [(0, 2, 1), (0, 1, 3), (1, 2, 3), (2, 0, 3)],
}


def write_exodus_3d(path):
    """A unit cube as Exodus II (NetCDF 64-bit offset, written by scipy):
    block 7 holds the 4 HEX8 of its lower half (2 x 2 x 1), block 9 the 24
    TET4 of its upper half (Kuhn split, 6 per cube); every element side on
    the cube's surface is in side set 11 + (face index), with Exodus side
    numbers."""
    from scipy.io import netcdf_file
    nx = 2
    g = np.linspace(0.0, 1.0, nx + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    nid = lambda i, j, k: (i * (nx + 1) + j) * (nx + 1) + k
    hexes, tets = [], []
    kuhn = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    for k in range(nx):
        for j in range(nx):
            for i in range(nx):
                h = [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)]
                if k == 0:
                    hexes.append(h)
                else:
                    tets += [[h[q] for q in t] for t in kuhn]
    elems = [("HEX8", e) for e in hexes] + [("TETRA", e) for e in tets]
    sides = {}
    for gi, (t, e) in enumerate(elems):
        for s, face in enumerate(EXO_FACES[t]):
            c = xyz[[e[q - 1] for q in face]]
            for ax in range(3):
                for side, val in ((0, 0.0), (1, 1.0)):
                    if np.all(c[:, ax] == val):
                        sides.setdefault(11 + 2 * ax + side, []).append((gi + 1, s + 1))
    f = netcdf_file(path, "w", version=2)
    f.title = b"cube"
    for name, n in (("len_string", 33), ("len_name", 33), ("four", 4), ("num_dim", 3), ("num_nodes", len(xyz)),
                    ("num_elem", len(elems)), ("num_el_blk", 2), ("num_side_sets", len(sides)),
                    ("num_el_in_blk1", len(hexes)), ("num_nod_per_el1", 8), ("num_el_in_blk2", len(tets)),
                    ("num_nod_per_el2", 4)):
        f.createDimension(name, n)
    for k, nm in enumerate(("coordx", "coordy", "coordz")):
        v = f.createVariable(nm, "d", ("num_nodes",))
        v[:] = xyz[:, k]
    v = f.createVariable("eb_prop1", "i", ("num_el_blk",))
    v[:] = [7, 9]
    v.name = b"ID"
    c1 = f.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
    c1[:] = np.array(hexes) + 1
    c1.elem_type = b"HEX8"
    c2 = f.createVariable("connect2", "i", ("num_el_in_blk2", "num_nod_per_el2"))
    c2[:] = np.array(tets) + 1
    c2.elem_type = b"TETRA"
    ids = sorted(sides)
    p = f.createVariable("ss_prop1", "i", ("num_side_sets",))
    p[:] = ids
    for k, sid in enumerate(ids, 1):
        f.createDimension(f"num_side_ss{k}", len(sides[sid]))
        e = f.createVariable(f"elem_ss{k}", "i", (f"num_side_ss{k}",))
        s = f.createVariable(f"side_ss{k}", "i", (f"num_side_ss{k}",))
        e[:] = [a for a, _ in sides[sid]]
        s[:] = [b for _, b in sides[sid]]
    f.close()
    return xyz, elems, sides


def test_exodus_3d_blocks_and_side_numbers(tmp_path):
    """A TET4 + HEX8 Exodus mesh written by scipy: nodes, connectivity and
    blocks come through, and every Exodus side number lands on the libMesh
    side with the same nodes."""
    mesh = tmp_path / "cube.e"
    xyz, elems, sides = write_exodus_3d(str(mesh))
    inp = open(os.path.join(INPUTS, "c1_hex8_reflective.i")).read()
    mb = inp[inp.index("[Mesh]"):inp.index("[]\n\n", inp.index("[Mesh]")) + 3]
    inp = inp.replace(mb, f"[Mesh]\n  [file]\n    type = FileMeshGenerator\n    file = '{mesh}'\n  []\n[]\n")
    inp = inp.replace('blocks = "0"', 'blocks = "7 9"').replace('materials = "1"', 'materials = "1 1"')
    inp = inp.replace("boundary = 'back front top right left bottom'", "boundary = '11 12 13 14 15 16'")
    (tmp_path / "cube.i").write_text(inp)
    d = tmp_path / "cube.dump"
    r = run(["-i", str(tmp_path / "cube.i"), "MonteCarlo/particles=200", "--dump-problem", str(d)])
    assert r.returncode == 0, r.stderr + r.stdout
    p = P.load_dump(str(d))
    m = p.mesh
    assert np.array_equal(m.xyz, xyz) and m.n_elem == len(elems)
    assert np.array_equal(m.subdomain, [7] * 4 + [9] * 24)
    off, fb = m.conn_offsets, m.face_boundary
    slot = np.concatenate([[0], np.cumsum([6 if t == "HEX8" else 4 for t, _ in elems])])
    for (t, e), c0 in zip(elems, off[:-1]):
        assert list(m.conn[c0:c0 + len(e)]) == list(e)
    seen = 0
    for sid, lst in sides.items():
        for gi, s in lst:
            t, e = elems[gi - 1]
            nodes = {e[q - 1] for q in EXO_FACES[t][s - 1]}
            lm = [k for k, f in enumerate(LM_FACES[t]) if {e[q] for q in f} == nodes]
            assert len(lm) == 1
            assert fb[slot[gi - 1] + lm[0]] == sid
            seen += 1
    assert seen == 12 + 24 and np.count_nonzero(fb >= 0) == seen  # 12 quads, 24 triangles on the surface
    o = __import__("oracle.pyoracle", fromlist=["Oracle"]).Oracle(p)
    st = o.run_batch(1, 200, threads=2)
    assert st["lost"] == 0 and st["leaks"] == 0


@pytest.mark.parametrize("content,message", [
    (b"\x89HDF\r\n\x1a\n" + b"\0" * 64, "NetCDF-4 (HDF5)"),
    (b"not a mesh at all", "not a NetCDF file"),
    (b"CDF\x02\0\0\0\0\0\0\0\x0a\0\0\0\x05", "truncated NetCDF header"),
])
def test_exodus_reader_errors(tmp_path, content, message):
    bad = tmp_path / "bad.e"
    bad.write_bytes(content)
    inp = open(os.path.join(INPUTS, "beavrs_assembly_3d.i")).read().replace("'31enr_16_in.e'", f"'{bad}'")
    (tmp_path / "bad.i").write_text(inp)
    r = run(["-i", str(tmp_path / "bad.i"), "--check-input"])
    assert r.returncode == 1 and message in r.stderr, r.stderr
