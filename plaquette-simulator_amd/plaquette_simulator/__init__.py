"""MI355X-native drop-in for plaquette_simulator (the reference's
plaquette_simulator/__init__.py:2): the same ``CliffordState`` class, backed
by the gfx950 bit-plane tableau engine."""
from .simulator import Batch, CliffordState, Program, SamplingState, Server
from .circuit import CliffordCircuit

__version__ = "0.0.1-alpha"
