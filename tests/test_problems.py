"""Synthetic problem generators: the slab-parallel Delaunay equals the single
triangulation, and the tet meshes tile the unit cube."""
import numpy as np

from macaw_amd import problems as P


def _lattice(n, seed):
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    h = 1.0 / (n - 1)
    j = rng.uniform(-0.25 * h, 0.25 * h, p.shape)
    j[(p == 0.0) | (p == 1.0)] = 0.0
    return p + j


def test_slab_delaunay_equals_single():
    for n, slabs, seed in ((13, 3, 1), (21, 5, 2)):
        x = _lattice(n, seed)
        a = P.delaunay_lattice(x, n, 0)
        b = P.delaunay_lattice(x, n, slabs, workers=1)
        assert np.array_equal(a, b), n


def test_tet_cube_tiles_unit_cube():
    m = P.tet_cube(18, seed=7, cache=False)
    x, t = m.xyz, m.conn.reshape(-1, 4)
    a = x[t[:, 1]] - x[t[:, 0]]
    b = x[t[:, 2]] - x[t[:, 0]]
    c = x[t[:, 3]] - x[t[:, 0]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    assert vol.min() > 0.0
    assert abs(vol.sum() - 1.0) < 1e-12
    # every interior face is shared by exactly two tets, boundary faces lie on the cube
    faces = np.sort(np.concatenate([t[:, [0, 1, 2]], t[:, [0, 1, 3]], t[:, [0, 2, 3]], t[:, [1, 2, 3]]]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2
    uf, cnt = np.unique(faces, axis=0, return_counts=True)
    bf = uf[cnt == 1]
    on = np.zeros(len(bf), bool)
    for d in range(3):
        for v in (0.0, 1.0):
            on |= np.all(x[bf][:, :, d] == v, axis=1)
    assert on.all()
