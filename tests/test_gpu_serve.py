"""The continuous-batching server: every request is served, and a request's result does not
depend on the other requests in the batch (the same steps and drag coefficient, bit for bit,
when it is served alone)."""
import pytest

from orrery_amd import serve

pytestmark = pytest.mark.gpu


def test_requests_are_served_and_independent_of_the_mix():
    reqs = serve.requests(5, rate=50.0, seed=3, dims=((48, 32, 32), (72, 40, 40)))
    server = serve.Server(chunk=500)
    served, wall = server.run(reqs)
    assert sorted(s.request.id for s in served) == [r.id for r in reqs]
    for s in served:
        print(f"measured: request {s.request.id} {s.request.body} {s.request.storage} steps {s.steps} "
              f"cd {s.cd:.6f} latency {s.latency:.3f} s", flush=True)
        assert s.finished >= s.admitted >= s.request.arrival
    probe = max(served, key=lambda s: s.request.cells)
    alone, _ = serve.Server(chunk=500).run([probe.request])
    assert (alone[0].steps, alone[0].cd, alone[0].converged) == (probe.steps, probe.cd, probe.converged)


def test_admission_waits_for_room_in_the_budget():
    reqs = serve.requests(3, rate=100.0, seed=4, dims=((48, 32, 32), (56, 36, 36)))
    from orrery_amd.batch import LatticeBatch
    need = max(LatticeBatch.required_bytes(*serve.request_lattice(r), r.storage) for r in reqs)
    server = serve.Server(budget_bytes=need + need // 4, chunk=500)   # room for one lattice at a time
    served, _ = server.run(reqs)
    assert len(served) == 3
    order = sorted(served, key=lambda s: s.admitted)
    for a, b in zip(order, order[1:]):
        assert b.admitted >= a.finished   # the next lattice entered only after the previous left
