"""ravenhip: Raven.jl's DGSEM right-hand side and LSRK time stepping on MI355X.

Host-side mirror of Raven.jl's API for this path over libravenhip.so (HIP
kernels for gfx950 behind the C ABI in include/ravenhip.h).
"""
from ._lib import RavenHIPError, lib  # noqa: F401
from .batch import Batch  # noqa: F401
from .dgsem import (DGSEM, AdvectionLaw, CentralFlux, EntropyConservativeFlux,  # noqa: F401
                    EulerLaw, FluxDifferencingForm, RusanovFlux, WeakForm, entropyintegral,
                    entropyproduct, weightednorm)
from .grids import (AdaptCoarsen, AdaptNone, AdaptRefine, AdaptTouched, GridManager, LobattoCell, brick, generate,  # noqa: F401
                    min_node_distance)
from .odesolvers import LSRK54, LSRK144, RLSRK54, RLSRK144, dostep_, solve_  # noqa: F401

lib()  # fail at import when the native library is missing
