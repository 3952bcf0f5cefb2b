"""stride_amd: MI355X-native iso-acoustic forward / adjoint solver with the
operator API of Stride's IsoAcousticDevito."""

from .problem import Space, Time, Grid, ScalarField, Traces, Shot, SubProblem  # noqa: F401
from .physics import IsoAcousticHIP, IsoAcousticDevito, ProblemTypeBase  # noqa: F401
from . import utils  # noqa: F401

__version__ = '0.1.0'
