"""readdygpu_amd: MI355X-native Brownian-dynamics step of ReaDDyGPU.

The compute path is libCReaDDyGPU.so (hand-written HIP for gfx950) behind the
C ABI in include/rdgpu.h; this package is its Python host mirror.
"""
from .model import (ATTRACTIVE, BOX, CYLINDER, DISK, HARMONIC, PIECEWISE_WELL, REPULSIVE, SPHERE, SPRING,
                    WEAK_HARMONIC, OrderOnePotential, OrderTwoPotential, ParticleType, SystemSpec)
from .simulation import RdError, Simulation

__all__ = [
    "ATTRACTIVE", "BOX", "CYLINDER", "DISK", "HARMONIC", "PIECEWISE_WELL", "REPULSIVE", "SPHERE", "SPRING",
    "WEAK_HARMONIC", "OrderOnePotential", "OrderTwoPotential", "ParticleType", "SystemSpec", "RdError",
    "Simulation",
]
