"""MI355X-native engine for OpenDust's ion-flow hot path.

The public classes are re-exported here so that both import styles of the
reference work: ``from opendust_amd import OpenDust`` and
``from opendust_amd.opendust import OpenDust``.  ``install_alias()`` makes
``opendust`` / ``opendust.opendust`` resolve to this package, so scripts
written for the reference run unchanged.
"""

import sys

from .opendust import (  # noqa: F401
    DustParticle,
    OpenDust,
    OutputParameters,
    PlasmaParametersInSIUnitsFieldDriven,
    PlasmaParametersInSIUnitsMaxwell,
    SimulatioParametersInSIUnits,
    SimulationParametersInSIUnits,
)

__all__ = [
    "DustParticle",
    "OpenDust",
    "OutputParameters",
    "PlasmaParametersInSIUnitsFieldDriven",
    "PlasmaParametersInSIUnitsMaxwell",
    "SimulatioParametersInSIUnits",
    "SimulationParametersInSIUnits",
    "install_alias",
]


def install_alias():
    """Register this package as ``opendust`` and ``opendust.opendust``."""
    from . import opendust as _mod

    sys.modules["opendust"] = sys.modules[__name__]
    sys.modules["opendust.opendust"] = _mod
