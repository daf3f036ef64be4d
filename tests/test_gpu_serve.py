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
    assert stats.completed == len(reqs) and stats.failed == 0
    assert stats.max_active >= 3  # several resident together (submission overlaps the ticks)
    server.close()
    cache = {}
    for r in reqs:
        assert np.array_equal(r.result, A.run_alone(r, cache)), (r.rid, r.kind)


def test_server_submit_poll_wait():
    """The scheduler's own protocol: submit, poll, wait, next_done; a request of zero
    steps returns its initial state; an invalid request is refused at submit."""
    import atum_hip as A
    from atum_hip import serve as S
    reqs = make_stream(duration=0.2, rate=60.0, seed=10, mix=(0.5, 0.2, 0.3), steps=(2, 6), vortex_k=(2, 3),
                       sw_k=(2, 3), imex_kh=(2, 3), imex_kv=3, nrange=(2, 3))
    server = A.Server(chunk=2, max_active=3)
    ids = [server.submit(r) for r in reqs]
    assert all(S._lib.lib.atm_server_poll(server.h, i) in (S.QUEUED, S.ACTIVE, S.DONE) for i in ids)
    assert S._lib.lib.atm_server_wait(server.h, ids[-1]) == S.DONE
    got = []
    while len(got) < len(reqs):
        r = server.next_done(10.0)
        assert r is not None and r.error is None
        assert r.finished >= r.admitted >= 0
        got.append(r)
    assert server.next_done(0.0) is None
    st = server.stats()
    assert st["max_active"] <= 3 and st["completed"] >= len(reqs)
    cache = {}
    for r in reqs:
        assert np.array_equal(r.result, A.run_alone(r, cache)), (r.rid, r.kind)
    # zero steps: the initial state comes back
    z = make_stream(duration=0.05, rate=200.0, seed=3, mix=(1.0, 0.0, 0.0), steps=(0, 0), vortex_k=(2, 2),
                    nrange=(2, 2))[0]
    server.submit(z)
    r = server.next_done(10.0)
    _, q0, _ = z.build({})
    assert r is z and np.array_equal(z.result, np.asarray(q0).reshape(z.result.shape))
    # a request whose state does not match its operator is refused before it is queued
    bad = make_stream(duration=0.05, rate=200.0, seed=3, mix=(1.0, 0.0, 0.0), steps=(1, 1), vortex_k=(2, 2),
                      nrange=(2, 2))[0]
    build = bad.build
    bad.build = lambda cache: (lambda dg, q0, dt: (dg, q0[:, :-1], dt))(*build(cache))
    with pytest.raises(ValueError):
        server.submit(bad)
    server.close()


def test_context_side_stream_tendency():
    """A tendency launched on a further stream of the context (atm_stream_create,
    atm_ctx_set_stream) equals the one on the context's own stream."""
    import atum_hip as A
    from gpu_cases import Case
    c = Case("euler", 2, 3, 3, sflux="rusanov", seed=4)
    q = c.dg.fieldarray(c.q)
    a, b = c.dg.fieldarray(), c.dg.fieldarray()
    c.dg(a, q, 0.0, increment=False)
    ctx = c.dg.ctx
    main = ctx.stream
    s = ctx.new_stream()
    ctx.stream_wait(s, main)
    ctx.use_stream(s)
    try:
        c.dg(b, q, 0.0, increment=False)
    finally:
        ctx.use_stream(main)
    ctx.stream_wait(main, s)
    ctx.sync()
    ctx.destroy_stream(s)
    np.testing.assert_array_equal(a.numpy(), b.numpy())


def test_server_refuses_invalid_requests():
    """Requests the scheduler cannot serve are refused at submit, before anything is
    queued: a footprint above the HBM budget, an operator of another context."""
    import atum_hip as A
    r = make_stream(duration=0.05, rate=200.0, seed=3, mix=(1.0, 0.0, 0.0), steps=(1, 1), vortex_k=(2, 2),
                    nrange=(2, 2))[0]
    small = A.Server(hbm_budget=1024.0)
    with pytest.raises(A.AtumError, match="budget"):
        small.submit(r)
    assert small.stats()["submitted"] == 0
    small.close()
    other = A.Server(ctx=A.Context(0))
    with pytest.raises(A.AtumError, match="another context"):
        other.submit(r)  # the request's operator lives on the default context
    other.close()
    ok = A.Server()
    ok.submit(r)
    done = ok.next_done(30.0)
    assert done is r and r.error is None
    ok.close()
