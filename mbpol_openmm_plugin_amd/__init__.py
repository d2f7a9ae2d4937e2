"""MI355X (gfx950) HIP platform for the MB-pol many-body water potential.

The product is ``libmbpol_hip.so`` (C ABI: ``include/mbpol_hip.h``); this
package holds its sources (``csrc/``), the force-field data (``params/``) and a
Python mirror of the reference plugin's Force / Context API (``api``).
"""
from .api import (Context, GuessModel, LangevinIntegrator, MBPolDispersionForce,  # noqa: F401
                  MBPolElectrostaticsForce, MBPolError, MBPolOneBodyForce, MBPolThreeBodyForce, MBPolTwoBodyForce,
                  State, System, ThreeParticleAverageSite, device_count, library)
from . import forcefield, systems  # noqa: F401
