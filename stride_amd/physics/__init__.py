"""Iso-acoustic physics: host-side preparation and the HIP-backed operator."""

from .iso_acoustic import ProblemTypeBase, IsoAcousticHIP, IsoAcousticDevito  # noqa: F401
from .plan import build_plan, adjoint_source_samples  # noqa: F401
