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


def _vortex_request(A, rid, dg, grid, steps=6):
    import problems as P
    from atum_hip.serve import Request
    q0 = P.isentropic_vortex(grid.points, 0.0, center=(0.0, 0.0), u0=(1.0, 1.0))
    return Request(rid, "vortex", ("vortex", rid), steps, 0.0, lambda cache: (dg, q0, 0.02))


def _vortex_op(A, K=4, N=3):
    import problems as P
    v = P.julia_range(-5, 5, K + 1)
    grid = A.brickgrid(A.LobattoCell(N, 2), (v, v), (True, True))
    dg = A.DGSEM(law=A.EulerLaw(2), grid=grid, volume_form=A.FluxDifferencingForm(A.EntropyConservativeFlux()),
                 surface_numericalflux=A.RusanovFlux(), precision="fast")
    return dg, grid


def _state_bytes(dg):
    grain = 1 << 16
    b = dg.Ns * dg.grid.Ntot * 8
    return 3 * ((b + grain - 1) // grain * grain)


def test_admission_counts_operator_memory():
    """The budget admits a request's buffers plus the device memory of its operators,
    an operator counted once while resident requests run it: with a budget that two
    requests' buffers fit but not their two operators as well, the requests on
    different operators run one at a time, the requests on one operator together;
    buffers of retired requests are reused from the pool."""
    import atum_hip as A
    L = A._lib.lib
    dga, grid = _vortex_op(A)
    dgb, _ = _vortex_op(A)
    sb = _state_bytes(dga)
    op = L.atm_dg_device_bytes(dga.h)
    assert op > 0 and op == L.atm_dg_device_bytes(dgb.h)
    budget = 2 * sb + 2 * op - 1  # both buffers fit; both operators as well do not
    assert 2 * sb <= budget
    # different operators: one at a time
    srv = A.Server(hbm_budget=float(budget), max_active=8, chunk=2)
    r1, r2 = _vortex_request(A, 1, dga, grid), _vortex_request(A, 2, dgb, grid)
    st = srv.run([r1, r2])
    assert st.completed == 2 and st.max_active == 1
    fp = srv.footprint_stats()
    assert fp["peak_bytes"] <= budget and fp["peak_operator_bytes"] == op
    assert fp["pool_hits"] >= 3  # the second request took the first one's three buffers
    srv.close()
    # one operator: counted once, so both are resident together
    srv = A.Server(hbm_budget=float(2 * sb + op), max_active=8, chunk=2)
    r3, r4 = _vortex_request(A, 3, dga, grid), _vortex_request(A, 4, dga, grid)
    st = srv.run([r3, r4])
    assert st.completed == 2 and st.max_active == 2
    assert srv.footprint_stats()["peak_bytes"] == 2 * sb + op
    srv.close()
    for r in (r1, r2, r3, r4):
        assert np.array_equal(r.result, A.run_alone(r))
    # buffers alone fit, buffers and operator do not: refused at submit
    small = A.Server(hbm_budget=float(sb + op - 1))
    with pytest.raises(A.AtumError, match="budget"):
        small.submit(_vortex_request(A, 5, dga, grid))
    small.close()
