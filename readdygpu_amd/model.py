"""Plain-data description of a Brownian-dynamics system.

Mirrors the model the reference builds in ``Simulation::parse``
(lib/ReaDDyGPU_Cpp/Simulation.cpp:206-280) from its XML input: particle types
with diffusion constants and radii (``ParticleType``, Simulation.hpp:19-28),
order-one and order-two potentials (Simulation.hpp:32-100) and the global
parameters (Simulation.hpp:138-181).  Nothing here computes; the GPU library
and the CPU oracle both consume it.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

# typeID / subtypeID numbering of Simulation.cpp:100-204
DISK, CYLINDER, SPHERE, BOX = 1, 2, 3, 4
HARMONIC, WEAK_HARMONIC, PIECEWISE_WELL = 1, 2, 3
ATTRACTIVE, REPULSIVE, SPRING = 1, 2, 3

_SUBTYPES = {"attractive": ATTRACTIVE, "repulsive": REPULSIVE, "spring": SPRING}


@dataclass
class ParticleType:
    """ParticleType (Simulation.hpp:19-28): name, D, default radius and the
    per-partner collision radii (0 -> default)."""
    name: str
    D: float
    radius: float
    collision_radii: Optional[Sequence[float]] = None


@dataclass
class OrderOnePotential:
    """DISK / CYLINDER / SPHERE / BOX (CudaOrderOnePotential, CudaSimulation.cuh:18-28)."""
    kind: int
    subtype: int
    k: float
    origin: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    normal: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    extension: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    height: float = 0.0
    radius: float = 0.0
    types: Optional[Sequence[int]] = None      # None -> all types


@dataclass
class OrderTwoPotential:
    """HARMONIC / WEAK_INTERACTION_HARMONIC / piecewise harmonic well
    (CudaOrderTwoPotential, CudaSimulation.cuh:30-37)."""
    kind: int
    subtype: int
    k: float
    length: float = 0.0
    depth: float = 0.0
    pairs: Optional[Sequence[Tuple[int, int]]] = None   # None -> all type pairs


@dataclass
class SystemSpec:
    """Global parameters (GlobalParametersHandler.cpp:73-143) + model."""
    box_lo: Tuple[float, float, float]
    box_hi: Tuple[float, float, float]
    dt: float
    kB: float = 0.0083144621
    T: float = 293.15
    max_cutoff: float = 10.0
    cell_size: float = 0.0          # <= 0: chosen from the interaction range
    seed: int = 1
    stream: int = 0
    types: List[ParticleType] = field(default_factory=list)
    pot1: List[OrderOnePotential] = field(default_factory=list)
    pot2: List[OrderTwoPotential] = field(default_factory=list)
    reaction_radii: Optional[np.ndarray] = None   # [nt, nt], 0 = no reaction
    # bonded groups (group potentials): pairs of global particle indices, the
    # potential of each bond, and per potential a rest length (None or < 0:
    # the collision radius sum of the pair's types)
    bonds: Optional[np.ndarray] = None            # [nb, 2]
    bond_pot: Optional[np.ndarray] = None         # [nb] index into bond_pots (None: all 0)
    bond_pots: List[OrderTwoPotential] = field(default_factory=list)
    bond_r0: Optional[Sequence[float]] = None

    @property
    def n_types(self) -> int:
        return len(self.types)

    @property
    def kbt(self) -> float:
        return self.kB * self.T

    def radius_sum(self) -> np.ndarray:
        """r0[i,j] = r_i(j) + r_j(i) (CudaSimulation.cu:357-391)."""
        nt = self.n_types
        r = np.zeros((nt, nt))
        for i, t in enumerate(self.types):
            for j in range(nt):
                v = 0.0 if t.collision_radii is None else float(t.collision_radii[j])
                r[i, j] = t.radius if v == 0.0 else v
        return r + r.T

    def pot2_mask(self, p: OrderTwoPotential) -> np.ndarray:
        nt = self.n_types
        m = np.zeros((nt, nt), dtype=np.uint8)
        if p.pairs is None:
            m[:] = 1
        else:
            for a, b in p.pairs:
                m[a, b] = 1
                m[b, a] = 1
        return m

    def pot1_mask(self, p: OrderOnePotential) -> int:
        if p.types is None:
            return (1 << self.n_types) - 1
        m = 0
        for t in p.types:
            m |= 1 << int(t)
        return m

    def pair_range(self) -> np.ndarray:
        """Range of the evaluated interaction per type pair: the largest natural
        range of the applicable potentials, capped at max_cutoff."""
        nt = self.n_types
        r0 = self.radius_sum()
        rc = np.zeros((nt, nt))
        for p in self.pot2:
            m = self.pot2_mask(p)
            if p.kind == HARMONIC:
                nat = r0 if p.subtype == REPULSIVE else np.full((nt, nt), self.max_cutoff)
            else:
                nat = r0 + p.length
            rc = np.where(m > 0, np.maximum(rc, np.minimum(nat, self.max_cutoff)), rc)
        return rc

    def rc2_f32(self) -> np.ndarray:
        rc = self.pair_range()
        return (rc * rc).astype(np.float32)

    def rr2_f32(self) -> np.ndarray:
        nt = self.n_types
        R = np.zeros((nt, nt)) if self.reaction_radii is None else np.maximum(
            self.reaction_radii, np.asarray(self.reaction_radii).T)
        return (R * R).astype(np.float32)

    def volume(self) -> float:
        return float(np.prod(np.asarray(self.box_hi, float) - np.asarray(self.box_lo, float)))


def subtype_id(name: str) -> int:
    return _SUBTYPES[name]
