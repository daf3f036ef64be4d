"""CPU tests of the product's host side: the C-ABI library loads and exports
every declared entry point, the host grid builder produces the same tables as
the oracle's (integer tables bit-exact), and solve!'s step logic matches."""
import os
import re

import numpy as np
import pytest

import problems as P
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            syms |= set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(atm_\w+)\s*\(", text, re.M))
    return syms


def test_library_exports_every_declared_symbol():
    import atum_hip._lib as L
    syms = _header_symbols()
    assert len(syms) > 20
    missing = [s for s in sorted(syms) if not hasattr(L.lib, s)]
    assert not missing, missing
    assert L.lib.atm_version().decode().startswith("atum-hip")


def test_no_device_needed_for_import():
    import atum_hip
    assert os.path.exists(atum_hip.LIB_PATH)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("ordering", ["cartesian", "stacked"])
@pytest.mark.parametrize("N", [1, 3, 4, 7])
def test_grid_tables_match_oracle(dim, ordering, N):
    import atum_hip as A
    coords = [P.julia_range(0, 2000, 4), P.julia_range(-5, 5, 3), P.julia_range(0, 1, 3)][:dim]
    per = [True, False, True][:dim]
    g = A.brickgrid(A.LobattoCell(N, dim), coords, per, ordering)
    o = O.Grid(N, coords, per, stacked=(ordering == "stacked"))
    for name in ("faceix_m", "faceix_p", "bctag"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    for name in ("points", "g", "MJ", "MJI", "faceMJ", "normal"):
        a, b = getattr(g, name), getattr(o, name)
        assert np.array_equal(a, b), (name, np.abs(a - b).max())
    assert np.allclose(g.D, o.Dmat, rtol=0, atol=4e-15 * np.abs(o.Dmat).max())
    ce, KV, nc = g.columns()
    oce, oKV, onc = o.colelem()
    assert KV == oKV and nc == onc and np.array_equal(ce, oce)


@pytest.mark.parametrize("N", range(1, 8))
def test_lgl_matches_oracle(N):
    from atum_hip.grid import lgl
    x, w, D = lgl(N)
    ox, ow, oD = O.lgl(N)
    assert np.allclose(x, ox, rtol=0, atol=2e-16)
    assert np.allclose(w, ow, rtol=4e-16)
    assert np.allclose(D, oD, rtol=0, atol=4e-15 * np.abs(oD).max())


class _FakeSolver:
    def __init__(self, dt):
        self.dt, self.time, self.steps = dt, 0.0, []

    def run(self, q, n):
        for _ in range(n):
            self.steps.append(self.dt)
            self.time += self.dt

    def dostep(self, q, after_stage=None):
        self.steps.append(self.dt)
        self.time += self.dt


@pytest.mark.parametrize("dt,timeend", [(0.125, 5.0), (0.0211289, 0.7), (104.6518, 432000.0),
                                        (0.1, 0.3)])
@pytest.mark.parametrize("adjust", [True, False])
def test_solve_step_logic(dt, timeend, adjust):
    """The batched fast path of solve_ takes the same steps as the reference loop
    (src/odesolvers.jl:9-26)."""
    from atum_hip.odesolvers import solve_
    fast = _FakeSolver(dt)
    n1 = solve_(None, timeend, fast, adjust_final=adjust)
    slow = _FakeSolver(dt)
    n2 = solve_(None, timeend, slow, after_step=lambda *a: None, adjust_final=adjust)
    assert n1 == n2 and fast.steps == slow.steps and fast.time == slow.time


def test_unsupported_pairs_rejected_without_gpu():
    """A law x flux pair the reference does not define is refused by the host
    mirror before any device work (Advection has no Roe flux)."""
    import atum_hip as A
    law = A.AdvectionLaw(2)
    assert not hasattr(law, "gamma")
    with pytest.raises(KeyError):
        A._lib.FLUX_IDS["upwind"]
