"""atum_hip: host-side mirror of Atum.jl's DGSEM / LSRK / ARK API over the
gfx950 C-ABI library libatumhip.so (include/atum_hip.h).

    law  = EulerLaw(2)
    cell = LobattoCell(4, 2)
    grid = brickgrid(cell, (vx, vy), periodic=(True, True))
    dg   = DGSEM(law=law, grid=grid, volume_form=FluxDifferencingForm(EntropyConservativeFlux()),
                 surface_numericalflux=RusanovFlux())
    q    = dg.fieldarray(initial_state)           # device field
    solve_(q, timeend, LSRK54(dg, q, dt))
"""
from ._lib import AtumError, Context, context, LIB_PATH  # noqa: F401  (fails loudly if unbuilt)
from .dgsem import DGSEM, FluxDifferencingForm, WeakForm  # noqa: F401
from .fields import Field, fieldarray  # noqa: F401
from .grid import Grid, LobattoCell, brickgrid, lgl  # noqa: F401
from .laws import (AdvectionLaw, CentralFlux, DummyProblem, EntropyConservativeFlux,  # noqa: F401
                   EulerGravityLaw, EulerLaw, KennedyGruberFlux, LinearEulerGravityLaw,
                   MatrixFlux, Problem, RoeFlux, RusanovFlux, ShallowWaterLaw)
from .measures import entropyintegral, entropyproduct, weightednorm  # noqa: F401
from .odesolvers import ARK23, LSRK54, LSRK144, RLSRK54, RLSRK144, BandedOperator, solve_  # noqa: F401
from .batch import Batch, LSRKEnsemble  # noqa: F401
from .serve import Request, Server, run_alone  # noqa: F401
from .snapshots import SnapshotWriter, load_checkpoint, read_snapshot, write_snapshot  # noqa: F401
from .tuning import TuneCache, autotune, autotune_banded, autotune_batch  # noqa: F401

__all__ = [n for n in dir() if not n.startswith("_")]
