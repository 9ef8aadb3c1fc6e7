"""Visit records of traced histories (trace_histories, batched record mode)
against the oracle's segment-by-segment walk of the same histories: one
record per element visit (segments grouped up to each non-scatter event),
same elements and ending events, consecutive visits joined at their shared
point, the last visit ending at the history's final position, energy and
seed, and straight visits as long as their one segment.  The records then
write to a RayTracingExodus-layout file."""
import numpy as np
import pytest

from macaw_amd import abi
from macaw_amd.lib import Context
from macaw_amd.ray_output import write_rays_exodus

from test_gpu_parity import c1_small, small_tet

pytestmark = pytest.mark.gpu

EV_SCATTER = 3


def visits_of(el, ev, le):
    out, n, length = [], 0, 0.0
    for e, v, ln in zip(el, ev, le):
        n += 1
        length += ln
        if v != EV_SCATTER:
            out.append((int(e), int(v), n, length))
            n, length = 0, 0.0
    return out


@pytest.mark.parametrize("make", [c1_small, small_tet], ids=["c1-hex-reflective", "tet-vacuum"])
def test_trace_matches_oracle_walk(make, tmp_path):
    from oracle.pyoracle import Oracle
    p = make()
    N, T, batch = p.n_particles, 40, 2
    c = Context(p, mode=abi.MODE_BATCHED, record_histories=1, trace_histories=T)
    st = c.run_batch(batch, N)
    tr = c.trace()
    h = c.histories(N)
    assert st["trace_records"] == tr.size > 0
    assert tr["history"].max() == T - 1
    o = Oracle(p)
    for i in range(T):
        r = tr[tr["history"] == i]
        vis = visits_of(*o.trace(batch, N, i))
        assert list(r["element"]) == [v[0] for v in vis], i
        assert list(r["event"]) == [v[1] for v in vis], i
        assert list(r["visit"]) == list(range(len(vis)))
        assert np.array_equal(r["start"][1:], r["end"][:-1])
        end = np.array([h["x"][i], h["y"][i], h["z"][i]])
        assert np.array_equal(r["end"][-1], end)
        assert r["energy"][-1] == h["E"][i] and r["seed_tracking"][-1] == h["seed"][i]
        for (e, ev, nseg, length), rr in zip(vis, r):
            if nseg == 1:
                # a segment of length L spans L |u|; |u| drifts from 1 by rounding over the
                # thousands of reflections of a reflective box (2.8e-12 seen)
                assert abs(np.linalg.norm(rr["end"] - rr["start"]) - length) <= 1e-10 * (1.0 + length)
    out = write_rays_exodus(str(tmp_path / "rays.e"), tr, batch=batch, n_particles=N)
    assert out == {"elements": tr.size, "nodes": tr.size + T, "rays": T}


def test_trace_needs_batched_record_mode():
    p = small_tet(n_particles=200)
    c = Context(p, mode=abi.MODE_PLAIN, record_histories=1, trace_histories=10)
    with pytest.raises(Exception, match="trace_histories"):
        c.run_batch(1, p.n_particles)
    c.set_options(mode=abi.MODE_BATCHED, record_histories=0)
    with pytest.raises(Exception, match="trace_histories"):
        c.run_batch(1, p.n_particles)
    c.set_options(record_histories=1, trace_histories=0)
    c.run_batch(1, p.n_particles)
    assert c.trace().size == 0
