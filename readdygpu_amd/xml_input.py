"""Reads the reference's six XML input files through the library's reader
(rd_model_read_xml) into a SystemSpec plus particle arrays.

Mirrors Simulation::readXML + readGlobalXML + parse
(lib/ReaDDyGPU_Cpp/Simulation.cpp:206-338), with the bonded pairs of the group
files (GroupParametersHandler.cpp, GroupReader.cpp:13-40) in
``spec.bonds``.  Parsing runs on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from .model import OrderOnePotential, OrderTwoPotential, ParticleType, SystemSpec

FILES = ("param_particles", "tplgy_potentials", "param_groups", "tplgy_coordinates", "tplgy_groups",
         "param_global")


@dataclass
class XmlInput:
    spec: SystemSpec
    xyz: np.ndarray
    types: np.ndarray
    readdy_ids: np.ndarray
    frames: int
    steps_per_frame: int
    n_groups: int
    device: int


def paths_in(directory: str) -> dict:
    return {k: os.path.join(directory, k + ".xml") for k in FILES}


def read_xml(paths: dict, max_cutoff: float = 10.0, seed: int = 1) -> XmlInput:
    lib = _lib.load()
    m = C.c_void_p()
    err = C.create_string_buffer(512)
    args = [paths[k].encode() for k in FILES]
    rc = lib.rd_model_read_xml(*args, C.byref(m), err, 512)
    if rc != 0:
        raise ValueError(err.value.decode())
    try:
        info = _lib.rd_model_info()
        lib.rd_model_get_info(m, C.byref(info))
        nt, n = info.n_types, info.n_particles
        D = np.zeros(nt)
        rad = np.zeros(nt)
        radii = np.zeros((nt, nt))
        lib.rd_model_types(m, D.ctypes.data_as(C.c_void_p), rad.ctypes.data_as(C.c_void_p),
                           radii.ctypes.data_as(C.c_void_p))
        xyz = np.zeros((n, 3), np.float32)
        types = np.zeros(n, np.int32)
        rid = np.zeros(n, np.int32)
        lib.rd_model_particles(m, xyz.ctypes.data_as(C.c_void_p), types.ctypes.data_as(C.c_void_p),
                               rid.ctypes.data_as(C.c_void_p))
        b = list(info.lattice_bounds)
        spec = SystemSpec(box_lo=(b[0], b[2], b[4]), box_hi=(b[1], b[3], b[5]), dt=info.dt_step,
                          kB=info.boltzmann, T=info.temperature, max_cutoff=max_cutoff, seed=seed)
        for i in range(nt):
            spec.types.append(ParticleType(f"type{i}", D=float(D[i]), radius=float(rad[i]),
                                           collision_radii=list(radii[i]) if radii[i].any() else None))
        for i in range(info.n_pot1):
            q = _lib.rd_pot1()
            lib.rd_model_pot1(m, i, C.byref(q))
            spec.pot1.append(OrderOnePotential(q.type, q.subtype, q.force_const, tuple(q.origin), tuple(q.normal),
                                               tuple(q.extension), q.height, q.radius,
                                               types=[t for t in range(nt) if (q.type_mask >> t) & 1]))
        for i in range(info.n_pot2):
            q = _lib.rd_pot2()
            mask = np.zeros((nt, nt), np.uint8)
            lib.rd_model_pot2(m, i, C.byref(q), mask.ctypes.data_as(C.c_void_p))
            pairs = [(a, c) for a in range(nt) for c in range(a, nt) if mask[a, c]]
            spec.pot2.append(OrderTwoPotential(q.type, q.subtype, q.force_const, q.length, q.depth, pairs=pairs))
        if info.n_bonds > 0:
            nb, npot = info.n_bonds, info.n_bond_pots
            bonds = np.zeros((nb, 2), np.int32)
            pot = np.zeros(nb, np.int32)
            pots = (_lib.rd_pot2 * npot)()
            r0 = np.zeros(npot)
            lib.rd_model_bonds(m, bonds.ctypes.data_as(C.c_void_p), pot.ctypes.data_as(C.c_void_p), pots,
                               r0.ctypes.data_as(C.c_void_p))
            spec.bonds, spec.bond_pot = bonds, pot
            spec.bond_pots = [OrderTwoPotential(q.type, q.subtype, q.force_const, q.length, q.depth) for q in pots]
            spec.bond_r0 = [None if v < 0 else float(v) for v in r0]
        return XmlInput(spec, xyz, types, rid, info.frames, info.steps_per_frame, info.n_groups, info.device)
    finally:
        lib.rd_model_free(m)
