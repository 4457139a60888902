"""Re-export of the compiled engine, as the reference's simulator.py:1 does.

Importing fails loudly when the HIP extension is not built: there is no CPU
fallback."""
from plaquette_simulator_bindings import Batch, CliffordState, Program, Server  # noqa: F401
from plaquette_simulator_bindings import sampling_state as SamplingState  # noqa: F401
