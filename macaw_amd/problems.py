"""Synthetic fixed-source problems: meshes, nuclides, materials, sources.

Every generator is seeded and deterministic, so the GPU library and the CPU
oracle are fed identical inputs.  A `Problem` is plain numpy data laid out
as the C ABI takes it (mcw_mesh_set, mcw_xs_set, ...).

Shapes follow the reference's inputs:
  * `hex_box` is MOOSE's GeneratedMeshGenerator (HEX8, libMesh node order,
    boundary ids 0 back, 1 bottom, 2 right, 3 top, 4 left, 5 front;
    test/tests/study/openmc_study.i:1-23);
  * `tet_cube` is a Delaunay triangulation of a jittered lattice;
  * nuclides are pointwise tables of 1/v absorption plus single-level
    Breit-Wigner resonances with constant potential scattering.
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import abi


@dataclasses.dataclass
class Mesh:
    xyz: np.ndarray            # (n_nodes, 3) f64
    elem_type: np.ndarray      # (n_elem,) u8
    conn_offsets: np.ndarray   # (n_elem+1,) i64
    conn: np.ndarray           # i32
    subdomain: np.ndarray      # (n_elem,) i32
    face_boundary: Optional[np.ndarray] = None  # per side slot i32

    @property
    def n_nodes(self):
        return self.xyz.shape[0]

    @property
    def n_elem(self):
        return self.elem_type.shape[0]


@dataclasses.dataclass
class NuclideData:
    awr: float
    energy: np.ndarray
    total: np.ndarray
    absorption: np.ndarray
    scatter: np.ndarray
    fission: Optional[np.ndarray] = None   # part of absorption; makes the nuclide fissionable
    nu: Tuple[float, float] = (0.0, 0.0)   # nu(E) = nu[0] + nu[1] E
    watt: Tuple[float, float] = (0.988e6, 2.249e-6)  # fission spectrum a [eV], b [1/eV]
    temperature: float = 0.0               # K of this table (0: 293.6)
    kappa: float = 0.0                     # energy released per fission [eV] (kappa-fission scores)


@dataclasses.dataclass
class MaterialData:
    id: int
    nuclides: np.ndarray   # i32
    densities: np.ndarray  # f64 [1/(b cm)]


@dataclasses.dataclass
class TallySpec:
    id: int
    estimator: int
    filters: List[dict]      # {"type": FILTER_*, "ids": [...]} or {"type": ENERGY, "edges": [...]}
    scores: List[int]
    nuclides: List[int] = dataclasses.field(default_factory=lambda: [-1])


@dataclasses.dataclass
class Problem:
    name: str
    mesh: Mesh
    nuclides: List[NuclideData]
    materials: List[MaterialData]
    block_material: Dict[int, int]
    bcs: Dict[int, int]
    source: abi.Source
    tallies: List[TallySpec]
    n_particles: int
    n_batches: int
    # further tables per nuclide index (other temperatures), element
    # temperatures [K] in mesh order, and the nearest-table tolerance [K]
    nuclide_temps: Dict[int, List[NuclideData]] = dataclasses.field(default_factory=dict)
    elem_temperature: Optional[np.ndarray] = None
    temperature_tolerance: float = 10.0


# ----------------------------------------------------------------- nuclides

def synth_nuclide(n_points: int, awr: float, seed: int, sigma_a0: float = 1.0,
                  sigma_p: float = 4.7, n_res: int = 50, e_min: float = 1e-5,
                  e_max: float = 2e7, jitter: float = 0.0, temperature: float = 0.0) -> NuclideData:
    """A pointwise nuclide: sigma_a = sigma_a0 sqrt(0.0253/E) + SLBW capture,
    sigma_s = sigma_p + SLBW elastic, on a log grid of n_points in [e_min, e_max] eV.

    Resonances: E_r log-uniform in [1, 1e5] eV, Gamma log-uniform in
    [1e-2, 10] eV, peak log-uniform in [10, 1e4] b, split Gamma_n/Gamma_gamma
    by a uniform fraction in [0.2, 0.8].  With jitter > 0 (below 0.5) every
    interior grid point moves by U(-jitter, jitter) log-steps, so nuclides of
    different seeds have different grids (a union grid of all their points).

    temperature > 0 [K] broadens every resonance for the Doppler width
    Gamma_D = sqrt(4 k T E_r / awr): width sqrt(Gamma^2 + Gamma_D^2), peak
    scaled to keep Gamma * peak (the same resonances at every temperature
    for one seed)."""
    rng = np.random.default_rng(seed)
    E = np.geomspace(e_min, e_max, n_points)
    E[0], E[-1] = e_min, e_max
    Er = np.exp(rng.uniform(np.log(1.0), np.log(1e5), n_res))
    G = np.exp(rng.uniform(np.log(1e-2), np.log(10.0), n_res))
    peak = np.exp(rng.uniform(np.log(10.0), np.log(1e4), n_res))
    fg = rng.uniform(0.2, 0.8, n_res)
    if jitter > 0.0:
        assert jitter < 0.5
        step = (np.log(e_max) - np.log(e_min)) / (n_points - 1)
        E[1:-1] = np.exp(np.log(E[1:-1]) + rng.uniform(-jitter, jitter, n_points - 2) * step)
    if temperature > 0.0:
        GD = np.sqrt(4.0 * abi.K_BOLTZMANN * temperature * Er / awr)
        Geff = np.sqrt(G * G + GD * GD)
        peak = peak * G / Geff
        G = Geff
    sa = sigma_a0 * np.sqrt(0.0253 / E)
    ss = np.full_like(E, sigma_p)
    for r in range(n_res):
        x = 2.0 * (E - Er[r]) / G[r]
        shape = peak[r] / (1.0 + x * x)
        sa += shape * fg[r] * np.sqrt(Er[r] / E)
        ss += shape * (1.0 - fg[r])
    st = sa + ss
    return NuclideData(awr=float(awr), energy=E, total=st, absorption=sa, scatter=ss,
                       temperature=float(temperature))


def flat_nuclide(awr: float, total: float, absorption: float, fission: float = 0.0,
                 nu=(0.0, 0.0), watt=(0.988e6, 2.249e-6), e_min: float = 1e-5,
                 e_max: float = 2e7, n_points: int = 2) -> NuclideData:
    """Energy-independent cross sections [b] on a log grid: the case where
    infinite-medium results have closed forms (k_inf = nu sigma_f / sigma_a)."""
    E = np.geomspace(e_min, e_max, n_points)
    E[0], E[-1] = e_min, e_max
    one = np.ones_like(E)
    return NuclideData(awr=float(awr), energy=E, total=total * one, absorption=absorption * one,
                       scatter=(total - absorption) * one,
                       fission=None if fission <= 0.0 else fission * one,
                       nu=(float(nu[0]), float(nu[1])), watt=(float(watt[0]), float(watt[1])))


def fissile(n: NuclideData, fraction: float, nu=(2.43, 0.0), watt=(0.988e6, 2.249e-6),
            kappa: float = 193.4e6) -> NuclideData:
    """The nuclide with a share `fraction` of its absorption made fission
    (kappa [eV] released per fission: U-235's 193.4 MeV by default)."""
    return dataclasses.replace(n, fission=fraction * np.asarray(n.absorption, np.float64),
                               nu=(float(nu[0]), float(nu[1])), watt=(float(watt[0]), float(watt[1])),
                               kappa=float(kappa))


# ----------------------------------------------------------------- meshes

HEX_SIDE_BOUNDARY = np.array([0, 1, 2, 3, 4, 5], np.int32)  # side s lies on boundary s


def hex_box(n: int, lo: float, hi: float) -> Mesh:
    """n^3 HEX8 box, GeneratedMeshGenerator layout (x fastest)."""
    m = n + 1
    g = np.linspace(lo, hi, m)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")          # x fastest
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()             # element i + n (j + n k)
    nid = lambda a, b, c: a + m * (b + m * c)
    conn = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                     nid(i, j + 1, k + 1)], 1).astype(np.int32)
    ne = n ** 3
    fb = np.tile(HEX_SIDE_BOUNDARY, ne)
    return Mesh(xyz=xyz, elem_type=np.full(ne, abi.HEX8, np.uint8),
                conn_offsets=np.arange(ne + 1, dtype=np.int64) * 8, conn=conn.ravel(),
                subdomain=np.zeros(ne, np.int32), face_boundary=fb)


def _cache_dir():
    d = os.environ.get("MACAW_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "macaw_amd"))
    os.makedirs(d, exist_ok=True)
    return d


def _delaunay_slab(args):
    """Tets of the Delaunay triangulation of `xyz` owned by one x-slab.

    The slab owns lattice layers [a, b) and sees `halo` more on each side.  A
    tet of the slab's local triangulation is kept when its centroid lies in
    the slab's x-range and its circumsphere holds no point of the whole set:
    the sphere is known to be empty when it stays between the nearest
    excluded layers, and is checked against every point (k-d tree) when it
    reaches past them.  For points in general position the Delaunay
    triangulation is unique, so the union over slabs is the global one."""
    from scipy.spatial import Delaunay, cKDTree
    xyz, layer, n, a, b, halo, S = args
    sel = np.nonzero((layer >= a - halo) & (layer < b + halo))[0]
    loc = Delaunay(xyz[sel]).simplices
    tets = sel[loc]
    p0 = xyz[tets[:, 0]]
    cen = xyz[tets].mean(axis=1)
    lo = -np.inf if a == 0 else a / (n - 1)
    hi = np.inf if b == n else b / (n - 1)
    own = (cen[:, 0] >= lo) & (cen[:, 0] < hi)
    tets, p0 = tets[own], p0[own]
    # circumcentre c solves 2 (p_k - p0) . c' = |p_k - p0|^2, c = p0 + c'
    A = xyz[tets[:, 1:]] - p0[:, None, :]
    rhs = 0.5 * np.einsum("nkd,nkd->nk", A, A)
    cc = np.linalg.solve(A, rhs[..., None])[..., 0]
    R = np.linalg.norm(cc, axis=1)
    c = p0 + cc
    ex_lo = xyz[layer < a - halo, 0].max() if a - halo > 0 else -np.inf
    ex_hi = xyz[layer >= b + halo, 0].min() if b + halo < n else np.inf
    safe = (c[:, 0] - R > ex_lo) & (c[:, 0] + R < ex_hi)
    keep = safe.copy()
    check = np.nonzero(~safe)[0]
    if check.size:
        tree = cKDTree(xyz)
        for i in check:
            inside = tree.query_ball_point(c[i], R[i] * (1.0 - 1e-9))
            keep[i] = all(j in set(tets[i]) for j in inside)
    return tets[keep]


def _canonical_tets(xyz, tets):
    """Sort each tet's node ids, order the tets lexicographically, then swap
    nodes 1 and 2 of negatively oriented tets so every volume is positive."""
    tets = np.sort(tets, axis=1)
    tets = tets[np.lexsort(tets.T[::-1])]
    a = xyz[tets[:, 1]] - xyz[tets[:, 0]]
    b = xyz[tets[:, 2]] - xyz[tets[:, 0]]
    c = xyz[tets[:, 3]] - xyz[tets[:, 0]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c))
    neg = vol < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return np.ascontiguousarray(tets, np.int32)


def delaunay_lattice(xyz: np.ndarray, n_lattice: int, n_slabs: int = 0, halo: int = 3, workers: int = 0):
    """Delaunay tets of a jittered n^3 lattice (point id = (i n + j) n + k, i
    the x layer), in canonical order.  `n_slabs` > 1 splits the work into
    x-slabs triangulated in parallel processes; the result is identical to
    the single triangulation."""
    n = n_lattice
    if n_slabs <= 1:
        from scipy.spatial import Delaunay
        return _canonical_tets(xyz, Delaunay(xyz).simplices)
    layer = np.arange(xyz.shape[0]) // (n * n)
    cuts = np.linspace(0, n, n_slabs + 1).round().astype(int)
    jobs = [(xyz, layer, n, int(cuts[s]), int(cuts[s + 1]), halo, s) for s in range(n_slabs)]
    workers = workers or min(n_slabs, os.cpu_count() or 1, 16)
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            parts = pool.map(_delaunay_slab, jobs)
            pool.close()
            pool.join()  # every worker reaped before the caller goes on
    else:
        parts = [_delaunay_slab(j) for j in jobs]
    return _canonical_tets(xyz, np.concatenate(parts))


def tet_cube(n_lattice: int, seed: int = 101, jitter: float = 0.25, cache: bool = True) -> Mesh:
    """Delaunay TET4 mesh of the unit cube from a jittered n^3 point lattice.

    Interior points move by up to +-jitter lattice spacings; points on a face,
    edge or corner stay on it.  Tets are in canonical order (sorted node ids,
    lexicographic) and oriented positively.  All external sides are boundary
    0.  Lattices above 40 points a side are triangulated in parallel x-slabs
    (identical result)."""
    key = hashlib.sha1(f"tet_cube-v2-{n_lattice}-{seed}-{jitter}".encode()).hexdigest()[:16]
    path = os.path.join(_cache_dir(), f"tet_cube_{n_lattice}_{key}.npz")
    if cache and os.path.exists(path):
        z = np.load(path)
        xyz, tets = z["xyz"], z["tets"]
    else:
        rng = np.random.default_rng(seed)
        g = np.linspace(0.0, 1.0, n_lattice)
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        h = 1.0 / (n_lattice - 1)
        J = rng.uniform(-jitter * h, jitter * h, P.shape)
        J[(P == 0.0) | (P == 1.0)] = 0.0
        xyz = P + J
        slabs = 0 if n_lattice <= 40 else max(4, n_lattice // 6)
        tets = delaunay_lattice(xyz, n_lattice, slabs)
        if cache:
            np.savez(path, xyz=xyz, tets=tets)
    ne = tets.shape[0]
    return Mesh(xyz=np.ascontiguousarray(xyz, np.float64), elem_type=np.full(ne, abi.TET4, np.uint8),
                conn_offsets=np.arange(ne + 1, dtype=np.int64) * 4,
                conn=np.ascontiguousarray(tets.ravel(), np.int32),
                subdomain=np.zeros(ne, np.int32), face_boundary=None)


def graded_mixed_box(n: int, m_tet: int, seed: int = 606, gamma=(1.4, 1.8)) -> Mesh:
    """Unit cube of n^3 tensor-product cells graded toward the corner
    (0, 0, 0): node i of axis d sits at (i/n)^gamma_d, gamma_d drawn from
    U(gamma) with `seed`.  Cells with i, j, k < m_tet (the refined corner) are
    each split into 6 Kuhn TET4 (the paths 000 -> 111 through the three axis
    orders, conforming between cells); the others are HEX8 in libMesh node
    order.  Hex and tet regions meet through split interfaces (a hex side
    covered by two tet sides).  Tets are oriented positively; all external
    sides are boundary 0."""
    rng = np.random.default_rng(seed)
    g = [np.linspace(0.0, 1.0, n + 1) ** rng.uniform(*gamma) for _ in range(3)]
    X, Y, Z = np.meshgrid(g[0], g[1], g[2], indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    m1 = n + 1
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()

    def node(di, dj, dk, sel):
        return ((I[sel] + di) * m1 + (J[sel] + dj)) * m1 + (K[sel] + dk)

    is_tet = (I < m_tet) & (J < m_tet) & (K < m_tet)
    hx = ~is_tet
    hexes = np.stack([node(0, 0, 0, hx), node(1, 0, 0, hx), node(1, 1, 0, hx), node(0, 1, 0, hx),
                      node(0, 0, 1, hx), node(1, 0, 1, hx), node(1, 1, 1, hx), node(0, 1, 1, hx)], 1)
    tets = []
    for a, b, c in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        step = np.zeros((4, 3), int)
        step[1, a] = 1
        step[2] = step[1]
        step[2, b] = 1
        step[3] = 1
        tets.append(np.stack([node(*step[q], is_tet) for q in range(4)], 1))
    tets = np.stack(tets, 1).reshape(-1, 4)
    v = xyz[tets]
    vol = np.einsum("ij,ij->i", v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))
    neg = vol < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    nh, nt = hexes.shape[0], tets.shape[0]
    conn = np.concatenate([hexes.ravel(), tets.ravel()]).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(np.r_[np.full(nh, 8), np.full(nt, 4)])]).astype(np.int64)
    etype = np.r_[np.full(nh, abi.HEX8, np.uint8), np.full(nt, abi.TET4, np.uint8)]
    return Mesh(xyz=np.ascontiguousarray(xyz, np.float64), elem_type=etype, conn_offsets=off, conn=conn,
                subdomain=np.zeros(nh + nt, np.int32), face_boundary=None)


def centroids(mesh: Mesh) -> np.ndarray:
    """Element centroids (mean of the nodes), any mix of TET4 and HEX8."""
    nn = np.diff(mesh.conn_offsets)
    cen = np.zeros((mesh.n_elem, 3))
    for k in np.unique(nn):
        sel = np.nonzero(nn == k)[0]
        idx = mesh.conn_offsets[sel][:, None] + np.arange(k)[None, :]
        cen[sel] = mesh.xyz[mesh.conn[idx]].mean(1)
    return cen


def voronoi_blocks(mesh: Mesh, n_regions: int, seed: int) -> np.ndarray:
    """Subdomain per element: index of the nearest of n_regions seeded centres
    to the element centroid (ties to the lower index)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    lo, hi = mesh.xyz.min(0), mesh.xyz.max(0)
    centres = rng.uniform(lo, hi, (n_regions, 3))
    _, idx = cKDTree(centres).query(centroids(mesh))
    return idx.astype(np.int32)


# ----------------------------------------------------------------- sources

def point_source(p, E0=2.0e6) -> abi.Source:
    s = abi.Source()
    s.space, s.angle, s.energy = abi.SPACE_POINT, abi.ANGLE_ISOTROPIC, abi.ENERGY_MONO
    s.point[:] = list(p)
    s.E0 = E0
    return s


def volume_watt_source(a=0.988e6, b=2.249e-6) -> abi.Source:
    s = abi.Source()
    s.space, s.angle, s.energy = abi.SPACE_MESH_VOLUME, abi.ANGLE_ISOTROPIC, abi.ENERGY_WATT
    s.watt_a, s.watt_b = a, b
    return s


def box_source(lo, hi, E0=2.0e6) -> abi.Source:
    s = abi.Source()
    s.space, s.angle, s.energy = abi.SPACE_BOX, abi.ANGLE_ISOTROPIC, abi.ENERGY_MONO
    s.box_lo[:] = list(lo)
    s.box_hi[:] = list(hi)
    s.E0 = E0
    return s


def linear_mu(src: abi.Source, axis, beta) -> abi.Source:
    """Make `src` anisotropic: p(mu) ~ 1 + beta mu, mu the cosine to `axis`
    (normalised here), azimuth uniform."""
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    src.angle = abi.ANGLE_LINEAR_MU
    src.axis[:] = list(a)
    src.beta = float(beta)
    return src


# ----------------------------------------------------------------- configs

def flux_tallies(both=True):
    t = [TallySpec(id=1, estimator=abi.EST_TRACKLENGTH,
                   filters=[{"type": abi.FILTER_CELL}], scores=[abi.SCORE_FLUX])]
    if both:
        t.append(TallySpec(id=2, estimator=abi.EST_COLLISION,
                           filters=[{"type": abi.FILTER_CELL}], scores=[abi.SCORE_FLUX]))
    return t


def c1_hex_reflective(n_particles=10_000, n_batches=10, n=8, n_points=20_000) -> Problem:
    """8x8x8 HEX8 reflective 10 cm cube, one A=12 nuclide at 0.08 /b-cm,
    isotropic 2 MeV point source at the centre, track-length + collision flux."""
    mesh = hex_box(n, -5.0, 5.0)
    nuc = synth_nuclide(n_points, 12.0, seed=202, sigma_a0=1.0, sigma_p=4.7)
    mat = MaterialData(id=1, nuclides=np.array([0], np.int32), densities=np.array([0.08]))
    return Problem(name="c1_hex8_reflective", mesh=mesh, nuclides=[nuc], materials=[mat],
                   block_material={0: 1}, bcs={b: abi.BC_REFLECT for b in range(6)},
                   source=point_source((0.0, 0.0, 0.0), 2.0e6), tallies=flux_tallies(),
                   n_particles=n_particles, n_batches=n_batches)


def eigen_box(sig_t=10.0, sig_a=3.0, sig_f=1.2, nu=2.5, awr=235.0, dens=0.05, n=4, half=1.0,
              reflective=True, n_particles=2000, n_batches=5) -> Problem:
    """A cube of one flat fissile nuclide.  Reflective: an infinite medium,
    k_inf = nu sigma_f / sigma_a; every history ends in absorption."""
    nuc = flat_nuclide(awr, sig_t, sig_a, sig_f, nu=(nu, 0.0))
    mat = MaterialData(id=1, nuclides=np.array([0], np.int32), densities=np.array([dens]))
    return Problem(name="eigen_box", mesh=hex_box(n, -half, half), nuclides=[nuc], materials=[mat],
                   block_material={0: 1},
                   bcs={b: (abi.BC_REFLECT if reflective else abi.BC_VACUUM) for b in range(6)},
                   source=point_source((0.0, 0.0, 0.0), 2.0e6), tallies=flux_tallies(),
                   n_particles=n_particles, n_batches=n_batches)


C2_NUCLIDES = [  # (A, sigma_p [b], sigma_a0 [b])
    (1.0, 20.0, 0.33),
    (12.0, 4.7, 0.0035),
    (16.0, 3.8, 0.0002),
    (56.0, 11.0, 2.56),
]


def c2_nuclides(n_points=50_000, n_nuclides=4, jitter=0.0, temperature=0.0):
    out = []
    for i in range(n_nuclides):
        A, sp, sa = C2_NUCLIDES[i % len(C2_NUCLIDES)]
        out.append(synth_nuclide(n_points, A, seed=1000 + i, sigma_a0=sa, sigma_p=sp, jitter=jitter,
                                 temperature=temperature))
    return out


def c2_materials(n_materials=4, n_nuclides=4, seed=303, dens=(0.01, 0.1)):
    rng = np.random.default_rng(seed + 1)
    mats = []
    for m in range(n_materials):
        nuc = np.arange(n_nuclides, dtype=np.int32)
        if n_nuclides > 4:
            nuc = np.sort(rng.choice(n_nuclides, size=4, replace=False)).astype(np.int32)
        mats.append(MaterialData(id=m + 1, nuclides=nuc,
                                 densities=rng.uniform(dens[0], dens[1], nuc.size)))
    return mats


def c2_tet_vacuum(n_lattice=54, n_particles=10_000_000, n_batches=1, n_points=50_000,
                  n_materials=4, n_nuclides=4, tallies=None, grid_jitter=0.0) -> Problem:
    """Unit cube, vacuum, Delaunay TET4 (54^3 lattice -> 1.0M tets), Voronoi
    materials from 4 nuclides, uniform volumetric Watt source."""
    mesh = tet_cube(n_lattice, seed=101)
    mesh.subdomain = voronoi_blocks(mesh, n_materials, seed=303)
    nucs = c2_nuclides(n_points, n_nuclides, jitter=grid_jitter)
    mats = c2_materials(n_materials, n_nuclides)
    return Problem(name=f"c2_tet4_vacuum_{mesh.n_elem}", mesh=mesh, nuclides=nucs, materials=mats,
                   block_material={b: b + 1 for b in range(n_materials)}, bcs={0: abi.BC_VACUUM},
                   source=volume_watt_source(),
                   tallies=tallies if tallies is not None else flux_tallies(both=False),
                   n_particles=n_particles, n_batches=n_batches)


def fissile_nuclide(n_points=50_000, seed=1235, jitter=0.0, temperature=0.0) -> NuclideData:
    """A synthetic heavy fissile nuclide: A = 235, 1/v absorption of 680 b at
    0.0253 eV plus SLBW resonances plus a flat 1.5 b (fast absorption), 85% of
    absorption is fission, nu = 2.43, Watt fission spectrum (a = 0.988 MeV,
    b = 2.249 /MeV)."""
    n = synth_nuclide(n_points, 235.0, seed=seed, sigma_a0=680.0, sigma_p=10.0, jitter=jitter,
                      temperature=temperature)
    n = dataclasses.replace(n, absorption=n.absorption + 1.5, total=n.total + 1.5)
    return fissile(n, 0.85, nu=(2.43, 0.0))


def k2_tet_eigen(n_lattice=54, n_particles=10_000_000, n_batches=5, n_points=50_000, dens=(0.5, 1.0),
                 fuel=0.4, n_materials=4) -> Problem:
    """C2's unit cube and Voronoi materials, made a critical-ish system for
    eigenvalue runs: the 4 C2 nuclides at densities U(dens) plus the fissile
    nuclide at `fuel` /b-cm in every material, vacuum boundary; batch 1
    starts from a uniform volumetric Watt source."""
    mesh = tet_cube(n_lattice, seed=101, cache=n_lattice > 20)
    mesh.subdomain = voronoi_blocks(mesh, n_materials, seed=303)
    nucs = c2_nuclides(n_points, 4) + [fissile_nuclide(n_points)]
    mats = c2_materials(n_materials, 4, dens=dens)
    for m in mats:
        m.nuclides = np.append(m.nuclides, np.int32(4)).astype(np.int32)
        m.densities = np.append(m.densities, fuel)
    return Problem(name=f"k2_tet4_eigen_{mesh.n_elem}", mesh=mesh, nuclides=nucs, materials=mats,
                   block_material={b: b + 1 for b in range(n_materials)}, bcs={0: abi.BC_VACUUM},
                   source=volume_watt_source(), tallies=flux_tallies(both=False),
                   n_particles=n_particles, n_batches=n_batches)


TEMPERATURES = (293.6, 600.0, 900.0, 1200.0)


def with_temperatures(p: Problem, temps=TEMPERATURES, t_range=(300.0, 1200.0), n_points=None,
                      tolerance=200.0, axis=0) -> Problem:
    """p (built from c2_nuclides, plus fissile_nuclide as a fifth nuclide when
    present) with every nuclide tabulated at each of `temps` (the first
    replaces p's own table; each further temperature has a grid 37 points
    longer, so tables of one nuclide differ in grid too) and element
    temperatures rising linearly along `axis` over t_range across the mesh."""
    n_points = n_points or p.nuclides[0].energy.size
    nn = len(p.nuclides)

    def tables(T, extra):
        out = c2_nuclides(n_points + extra, min(nn, 4), temperature=T)
        if nn > 4:
            out.append(fissile_nuclide(n_points + extra, temperature=T))
        return out

    base = tables(temps[0], 0)
    more = {q: [] for q in range(nn)}
    for k, T in enumerate(temps[1:], 1):
        for q, t in enumerate(tables(T, 37 * k)):
            more[q].append(t)
    c = centroids(p.mesh)[:, axis]
    lo, hi = c.min(), c.max()
    T = t_range[0] + (t_range[1] - t_range[0]) * (c - lo) / max(hi - lo, 1e-300)
    return dataclasses.replace(p, name=p.name + "_temps", nuclides=base, nuclide_temps=more,
                               elem_temperature=T, temperature_tolerance=float(tolerance))


def c3_tet_vacuum(n_lattice=107, n_particles=100_000_000, n_batches=1) -> Problem:
    """C2's generator at 8M tets (107^3 lattice -> 8.07M tets) with 16
    Voronoi materials mixing 4 of 32 synthetic nuclides (50k-point grids,
    each jittered, so the device's union grid holds 1.6M energies)."""
    return c2_tet_vacuum(n_lattice=n_lattice, n_particles=n_particles, n_batches=n_batches,
                         n_materials=16, n_nuclides=32, grid_jitter=0.4)


C5_MESH = (190, 138)   # 19,999,360 elements: 15,768,432 TET4 + 4,230,928 HEX8


def c5_mixed(n=C5_MESH[0], m_tet=C5_MESH[1], n_particles=10_000_000, n_batches=1, n_points=100_000,
             n_nuclides=64, n_materials=128, seed=707) -> Problem:
    """Stress case: graded mixed TET4/HEX8 unit cube (190^3 cells, the
    138^3-cell refined corner as Kuhn tets: 20.0M elements), vacuum; 128
    Voronoi materials mixing 4 of 64 nuclides (100k-point jittered grids: a
    6.4M-point union) at low densities U(0.001, 0.01) (long tracks); uniform
    volumetric Watt source with directions linear in mu about a seeded axis
    (beta seeded in [0.3, 0.9])."""
    mesh = graded_mixed_box(n, m_tet, seed=606)
    mesh.subdomain = voronoi_blocks(mesh, n_materials, seed=seed)
    nucs = c2_nuclides(n_points, n_nuclides, jitter=0.4)
    mats = c2_materials(n_materials, n_nuclides, seed=seed, dens=(0.001, 0.01))
    rng = np.random.default_rng(seed + 2)
    src = linear_mu(volume_watt_source(), rng.normal(size=3), rng.uniform(0.3, 0.9))
    return Problem(name=f"c5_mixed_{mesh.n_elem}", mesh=mesh, nuclides=nucs, materials=mats,
                   block_material={b: b + 1 for b in range(n_materials)}, bcs={0: abi.BC_VACUUM},
                   source=src, tallies=flux_tallies(both=False), n_particles=n_particles, n_batches=n_batches)


C4_LATTICES = (26, 54, 85)   # 105,966 / 1,008,162 / 4,014,898 tets


def c4_meshes(lattices=C4_LATTICES, n_points=50_000) -> List[Problem]:
    """The serving mix's resident meshes: the C2 generator (unit cube, vacuum,
    4 Voronoi materials of the 4 C2 nuclides) at 100k, 1M and 4M tets.  Their
    own source and tallies are unused by the request server."""
    out = []
    for n in lattices:
        p = c2_tet_vacuum(n_lattice=n, n_particles=1, n_points=n_points, tallies=[])
        out.append(p)
    return out


@dataclasses.dataclass
class Request:
    request_id: int
    mesh_slot: int
    n_particles: int
    source: abi.Source
    arrival_s: float = 0.0


def c4_requests(n=2000, seed=505, n_meshes=3, n_min=1e4, n_max=1e7, rate=None) -> List[Request]:
    """Fixed-source requests: mesh uniform over the slots, particle count
    log-uniform in [n_min, n_max], a point (uniform in [0.05, 0.95]^3) or a
    box (corner uniform, sides uniform in [0.05, 0.3], inside the unit cube)
    source with equal odds, Watt energies.  With `rate` (requests/s) the
    arrival times are a Poisson process from the same seed."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        slot = int(rng.integers(n_meshes))
        npart = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        if rng.random() < 0.5:
            src = point_source(rng.uniform(0.05, 0.95, 3))
        else:
            side = rng.uniform(0.05, 0.3, 3)
            lo = rng.uniform(0.0, 1.0 - side)
            src = box_source(lo, lo + side)
        src.energy = abi.ENERGY_WATT
        src.watt_a, src.watt_b = 0.988e6, 2.249e-6
        out.append(Request(i, slot, npart, src))
    if rate:
        t = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, n))
        for r, a in zip(out, t):
            r.arrival_s = float(a)
    return out


# ----------------------------------------------------------------- marshalling

def nuclide_struct(n: NuclideData, keep: list) -> abi.Nuclide:
    """mcw_nuclide view of n; the arrays it points to are appended to keep."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (n.energy, n.total, n.absorption, n.scatter)]
    fis = None if n.fission is None else np.ascontiguousarray(n.fission, np.float64)
    keep += arrs + [fis]
    return abi.Nuclide(n.energy.size, n.awr, *[abi.ptr(a, abi.C.c_double) for a in arrs],
                       abi.ptr(fis, abi.C.c_double), (abi.C.c_double * 2)(*n.nu),
                       (abi.C.c_double * 2)(*n.watt), float(n.temperature), float(n.kappa))


class Marshal:
    """Holds ctypes views of a Problem's arrays alive while a library copies them."""

    def __init__(self, p: Problem):
        self.p = p
        self.keep = []
        self.nuclides = (abi.Nuclide * len(p.nuclides))()
        for i, n in enumerate(p.nuclides):
            self.nuclides[i] = self.nuclide(n)
        self.temps = [(int(q), self.nuclide(t)) for q, ts in sorted(p.nuclide_temps.items()) for t in ts]
        self.elem_T = (None if p.elem_temperature is None
                       else np.ascontiguousarray(p.elem_temperature, np.float64))
        self.materials = (abi.Material * len(p.materials))()
        for i, m in enumerate(p.materials):
            nu = np.ascontiguousarray(m.nuclides, np.int32)
            de = np.ascontiguousarray(m.densities, np.float64)
            self.keep += [nu, de]
            self.materials[i] = abi.Material(m.id, nu.size, abi.ptr(nu, abi.C.c_int32),
                                             abi.ptr(de, abi.C.c_double))
        self.blocks = np.array(list(p.block_material.keys()), np.int32)
        self.block_mats = np.array(list(p.block_material.values()), np.int32)

    def nuclide(self, n: NuclideData) -> abi.Nuclide:
        return nuclide_struct(n, self.keep)

    def filters(self, t: TallySpec):
        arr = (abi.Filter * max(1, len(t.filters)))()
        for k, f in enumerate(t.filters):
            ids = f.get("ids")
            edges = f.get("edges")
            ida = None if ids is None else np.ascontiguousarray(ids, np.int64)
            eda = None if edges is None else np.ascontiguousarray(edges, np.float64)
            self.keep += [ida, eda]
            n = len(ids) if ids is not None else (len(edges) if edges is not None else 0)
            arr[k] = abi.Filter(f["type"], n, abi.ptr(ida, abi.C.c_int64), abi.ptr(eda, abi.C.c_double))
        scores = np.ascontiguousarray(t.scores, np.int32)
        nucs = np.ascontiguousarray(t.nuclides, np.int32)
        self.keep += [scores, nucs]
        return arr, scores, nucs


def load_into(lib_prefix_calls, p: Problem):
    """Feed a Problem through a set of setter callables with the mcw_* signatures.

    `lib_prefix_calls` maps names ('mesh_set', 'bc_set', 'xs_set', 'xs_add_temperature',
    'temperature_set', 'materials_set', 'block_material_map', 'source_set', 'tally_add') to
    callables returning 0 on success and raising/returning non-zero on error."""
    m = p.mesh
    M = Marshal(p)
    c = lib_prefix_calls
    fb = None if m.face_boundary is None else np.ascontiguousarray(m.face_boundary, np.int32)
    c["mesh_set"](m.n_nodes, abi.ptr(np.ascontiguousarray(m.xyz, np.float64), abi.C.c_double), m.n_elem,
                  abi.ptr(m.elem_type, abi.C.c_uint8), abi.ptr(m.conn_offsets, abi.C.c_int64),
                  abi.ptr(m.conn, abi.C.c_int32), abi.ptr(m.subdomain, abi.C.c_int32),
                  abi.ptr(fb, abi.C.c_int32))
    for b, k in p.bcs.items():
        c["bc_set"](int(b), int(k))
    c["xs_set"](len(p.nuclides), M.nuclides)
    for q, t in M.temps:
        c["xs_add_temperature"](q, abi.C.byref(t))
    if M.elem_T is not None:
        c["temperature_set"](M.elem_T.size, abi.ptr(M.elem_T, abi.C.c_double), float(p.temperature_tolerance))
    c["materials_set"](len(p.materials), M.materials)
    c["block_material_map"](len(M.blocks), abi.ptr(M.blocks, abi.C.c_int32), abi.ptr(M.block_mats, abi.C.c_int32))
    c["source_set"](abi.C.byref(p.source))
    for t in p.tallies:
        f, s, n = M.filters(t)
        c["tally_add"](t.id, t.estimator, len(t.filters), f, len(t.scores), abi.ptr(s, abi.C.c_int32),
                       len(t.nuclides), abi.ptr(n, abi.C.c_int32))
    return M


# ----------------------------------------------------------------- driver dumps

def load_dump(path) -> Problem:
    """Problem written by `macaw-amd-opt --dump-problem` (the exact bytes the
    driver handed to the library), for replay on a CPU reference."""
    raw = open(path, "rb").read()
    versions = (b"MCWD0001", b"MCWD0002", b"MCWD0003", b"MCWD0004", b"MCWD0005")
    if raw[:8] not in versions:
        raise ValueError(f"{path}: not a problem dump")
    ver = versions.index(raw[:8]) + 1
    v2, v3, v4, v5 = ver >= 2, ver >= 3, ver >= 4, ver >= 5
    pos = [8]

    def i64():
        v = int(np.frombuffer(raw, "<i8", 1, pos[0])[0])
        pos[0] += 8
        return v

    def arr(dtype, n):
        a = np.frombuffer(raw, dtype, n, pos[0]).copy()
        pos[0] += a.nbytes
        return a

    nn = i64()
    xyz = arr("<f8", 3 * nn).reshape(nn, 3)
    ne = i64()
    etype = arr("u1", ne)
    off = arr("<i8", ne + 1)
    nc = i64()
    conn = arr("<i4", nc)
    sub = arr("<i4", ne)
    nfb = i64()
    fb = arr("<i4", nfb)
    bcs = {}
    for _ in range(i64()):
        b = i64()
        bcs[b] = i64()
    bm = {}
    for _ in range(i64()):
        b = i64()
        bm[b] = i64()
    def table():
        awr = float(arr("<f8", 1)[0])
        npt = i64()
        E, t, a, s = (arr("<f8", npt) for _ in range(4))
        n = NuclideData(awr=awr, energy=E, total=t, absorption=a, scatter=s)
        if v2:
            fis = arr("<f8", npt) if i64() else None
            nu, watt = arr("<f8", 2), arr("<f8", 2)
            n = dataclasses.replace(n, fission=fis, nu=(float(nu[0]), float(nu[1])),
                                    watt=(float(watt[0]), float(watt[1])))
        if v3:
            n = dataclasses.replace(n, temperature=float(arr("<f8", 1)[0]))
        if v4:
            n = dataclasses.replace(n, kappa=float(arr("<f8", 1)[0]))
        return n

    nucs, temps = [], {}
    for q in range(i64()):
        nucs.append(table())
        if v3:
            more = [table() for _ in range(i64())]
            if more:
                temps[q] = more
    mats = []
    for _ in range(i64()):
        mid = i64()
        k = i64()
        mats.append(MaterialData(id=mid, nuclides=arr("<i4", k), densities=arr("<f8", k)))
    src = abi.Source.from_buffer_copy(raw[pos[0]:pos[0] + abi.C.sizeof(abi.Source)])
    pos[0] += abi.C.sizeof(abi.Source)
    particles, batches, roundtrip, seed = i64(), i64(), i64(), i64()
    run_mode, n_inactive = (i64(), i64()) if v2 else (abi.RUN_FIXED_SOURCE, 0)
    elem_T, tol = None, 10.0
    if v3:
        k = i64()
        elem_T = arr("<f8", k) if k else None
        tol = float(arr("<f8", 1)[0])
    material_index = i64() if v5 else abi.MATERIAL_BY_ID
    mesh = Mesh(xyz=np.ascontiguousarray(xyz), elem_type=etype, conn_offsets=off, conn=conn, subdomain=sub,
                face_boundary=fb if nfb else None)
    p = Problem(name=os.path.basename(path), mesh=mesh, nuclides=nucs, materials=mats, block_material=bm,
                bcs=bcs, source=src, tallies=[], n_particles=particles, n_batches=batches,
                nuclide_temps=temps, elem_temperature=elem_T, temperature_tolerance=tol)
    p.seed_roundtrip = bool(roundtrip)
    p.master_seed = seed
    p.run_mode = run_mode
    p.n_inactive = n_inactive
    p.material_index = material_index
    return p
