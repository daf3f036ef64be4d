"""Serving loop: continuous batching with every result bitwise equal to the request alone."""
import numpy as np
import pytest

from serve_requests import make_stream

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    import atum_hip as A
    reqs = make_stream(duration=0.3, rate=60.0, seed=5, steps=(3, 17), vortex_k=(2, 6), sw_k=(2, 6),
                       imex_kh=(2, 3), imex_kv=3, nrange=(2, 5))
    assert len(reqs) >= 8 and {r.kind for r in reqs} == {"vortex", "sw", "imex"}
    server = A.Server(chunk=4)
    stats = server.run(reqs)
    s = stats.summary()
    assert s["completed"] == len(reqs)
    assert stats.max_active >= 2  # requests overlapped on the device
    assert s["p99_s"] >= s["p50_s"] > 0
    cache = {}
    for r in reqs:
        alone = A.run_alone(r, cache)
        assert np.array_equal(r.result, alone), (r.rid, r.kind)


def test_implicit_requests_on_side_streams():
    """IMEX requests spread over three streams beside the explicit batch: every
    result is still bitwise the request's own, with the streams joined each tick."""
    import atum_hip as A
    reqs = make_stream(duration=0.2, rate=60.0, seed=11, mix=(0.3, 0.0, 0.7), steps=(4, 13), vortex_k=(2, 4),
                       imex_kh=(2, 4), imex_kv=3, nrange=(3, 3))
    assert sum(r.kind == "imex" for r in reqs) >= 4
    for r in reqs:
        r.arrival = 0.0  # all admitted in the first tick: the side streams run together
    server = A.Server(chunk=3, imex_streams=3)
    stats = server.run(reqs)
    assert stats.completed == len(reqs)
    assert len(server._streams) == 3
    server.close()
    cache = {}
    for r in reqs:
        assert np.array_equal(r.result, A.run_alone(r, cache)), (r.rid, r.kind)
