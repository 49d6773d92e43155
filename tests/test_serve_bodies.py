"""The served bodies at the smallest and largest request sizes: closed, outward meshes of the
requested length along x (host only)."""
import numpy as np
import pytest

from orrery_amd import serve


@pytest.mark.parametrize("body", serve.BODIES)
@pytest.mark.parametrize("length", [8, 64])
def test_bodies_at_every_request_size(body, length):
    m = serve.body_mesh(body, float(length))
    t = m.triangles.astype(np.int64)
    edges = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    directed = {tuple(e) for e in edges.tolist()}
    assert len(directed) == len(edges) and all((b, a) in directed for a, b in directed)
    assert m.volume() > 0.0
    lo, hi = m.bounds()
    assert max(hi - lo) <= 1.5 * length + 1e-9
