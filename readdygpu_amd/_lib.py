"""ctypes bindings of libCReaDDyGPU.so (include/rdgpu.h).

The library is built in-tree by ``readdygpu_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback for the GPU path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libCReaDDyGPU.so")

RD_MAX_TYPES = 16


class rd_params(C.Structure):
    _fields_ = [
        ("lattice_bounds", C.c_double * 6),
        ("dt", C.c_double),
        ("boltzmann", C.c_double),
        ("temperature", C.c_double),
        ("max_cutoff", C.c_double),
        ("cell_size", C.c_double),
        ("seed", C.c_uint64),
        ("stream", C.c_uint32),
        ("device", C.c_int),
    ]


class rd_pot1(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("subtype", C.c_int),
        ("force_const", C.c_float),
        ("origin", C.c_float * 3),
        ("normal", C.c_float * 3),
        ("extension", C.c_float * 3),
        ("height", C.c_float),
        ("radius", C.c_float),
        ("type_mask", C.c_uint32),
    ]


class rd_pot2(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("subtype", C.c_int),
        ("force_const", C.c_float),
        ("length", C.c_float),
        ("depth", C.c_float),
    ]


class rd_launch_config(C.Structure):
    _fields_ = [
        ("kernel", C.c_int),
        ("used_kernel", C.c_int),
        ("block", C.c_int),
        ("used_block", C.c_int),
        ("ppt", C.c_int),
        ("used_ppt", C.c_int),
    ]


class rd_model_info(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int),
        ("n_types", C.c_int),
        ("n_pot1", C.c_int),
        ("n_pot2", C.c_int),
        ("n_groups", C.c_int),
        ("n_bonds", C.c_int),
        ("n_bond_pots", C.c_int),
        ("frames", C.c_int),
        ("steps_per_frame", C.c_int),
        ("dt_frame", C.c_double),
        ("dt_step", C.c_double),
        ("temperature", C.c_double),
        ("boltzmann", C.c_double),
        ("lattice_bounds", C.c_double * 6),
        ("device", C.c_int),
        ("testmode", C.c_int),
    ]


class rd_info(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("n_types", C.c_int),
        ("lattice", C.c_int * 3),
        ("cell_size", C.c_double),
        ("stencil", C.c_int),
        ("step", C.c_uint64),
        ("rdf_frames", C.c_long),
        ("n_replicas", C.c_int),
        ("n_cells", C.c_int),
    ]


class rd_train_params(C.Structure):
    _fields_ = [
        ("batch", C.c_int),
        ("steps", C.c_int),
        ("lr", C.c_float),
        ("lr_min", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("seed", C.c_uint64),
    ]


P = C.c_void_p
_SIGS = {
    "rd_create": (C.c_int, [C.POINTER(rd_params), C.POINTER(P)]),
    "rd_destroy": (None, [P]),
    "rd_last_error": (C.c_char_p, [P]),
    "rd_set_particle_types": (C.c_int, [P, C.c_int, P, P, P]),
    "rd_add_order_one_potential": (C.c_int, [P, C.POINTER(rd_pot1)]),
    "rd_add_order_two_potential": (C.c_int, [P, C.POINTER(rd_pot2), P]),
    "rd_set_reaction_radii": (C.c_int, [P, P]),
    "rd_set_bonds": (C.c_int, [P, C.c_int, P, P, C.c_int, C.POINTER(rd_pot2), P]),
    "rd_traj_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(P)]),
    "rd_traj_write_frame": (C.c_int, [P, C.c_int, C.c_int, P, P]),
    "rd_traj_close": (C.c_int, [P]),
    "rd_rdf_write": (C.c_int, [C.c_char_p, C.c_int, P, C.c_int, C.c_double, P]),
    "rd_apply_reactions_replicas": (C.c_int, [P, P, P]),
    "rd_simulate_async": (C.c_int, [P, C.c_int]),
    "rd_checkpoint_write": (C.c_int, [P, C.c_char_p]),
    "rd_checkpoint_read": (C.c_int, [P, C.c_char_p]),
    "rd_set_step": (C.c_int, [P, C.c_uint64]),
    "rd_check_invariants": (C.c_int, [P, C.POINTER(C.c_long)]),
    "rd_frame_copy": (C.c_int, [P]),
    "rd_frame_wait": (C.c_int, [P, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)]),
    "rd_frame_rewind": (C.c_int, [P]),
    "rd_set_particles": (C.c_int, [P, C.c_int, P, P]),
    "rd_set_replicas": (C.c_int, [P, C.c_int, P, P, P, P, P]),
    "rd_initialize": (C.c_int, [P]),
    "rd_simulate": (C.c_int, [P, C.c_int]),
    "rd_copy_pos_from_device": (C.c_int, [P, P, P]),
    "rd_cell_ids": (C.c_int, [P, P]),
    "rd_gradients": (C.c_int, [P, P, C.POINTER(C.c_double)]),
    "rd_noise_fill": (C.c_int, [P, C.c_uint64, C.c_int, P]),
    "rd_set_noise_array": (C.c_int, [P, P, C.c_int]),
    "rd_rdf_configure": (C.c_int, [P, C.c_double, C.c_int]),
    "rd_rdf_calculate": (C.c_int, [P]),
    "rd_rdf_copy": (C.c_int, [P, P, C.POINTER(C.c_long)]),
    "rd_rdf_normalize": (C.c_int, [P, P]),
    "rd_rdf_reset": (C.c_int, [P]),
    "rd_neighbor_pairs": (C.c_int, [P, P, C.c_long, C.POINTER(C.c_long)]),
    "rd_reaction_candidates": (C.c_int, [P, P, C.c_long, C.POINTER(C.c_long)]),
    "rd_reaction_candidates_noalloc": (C.c_int, [P, P, C.c_long, C.POINTER(C.c_long)]),
    "rd_apply_reactions": (C.c_int, [P, P]),
    "rd_get_info": (C.c_int, [P, C.POINTER(rd_info)]),
    "rd_set_launch_config": (C.c_int, [P, C.POINTER(rd_launch_config)]),
    "rd_get_launch_config": (C.c_int, [P, C.POINTER(rd_launch_config)]),
    "rd_autotune": (C.c_int, [P, C.c_int, C.POINTER(C.c_double)]),
    "rd_set_cell_size": (C.c_int, [P, C.c_double]),
    "rd_device_name": (C.c_int, [C.c_int, C.c_char_p, C.c_int]),
    "rd_model_read_xml": (C.c_int, [C.c_char_p] * 6 + [C.POINTER(P), C.c_char_p, C.c_int]),
    "rd_model_free": (None, [P]),
    "rd_model_get_info": (C.c_int, [P, C.POINTER(rd_model_info)]),
    "rd_model_types": (C.c_int, [P, P, P, P]),
    "rd_model_particles": (C.c_int, [P, P, P, P]),
    "rd_model_pot1": (C.c_int, [P, C.c_int, C.POINTER(rd_pot1)]),
    "rd_model_pot2": (C.c_int, [P, C.c_int, C.POINTER(rd_pot2), P]),
    "rd_model_bonds": (C.c_int, [P, P, P, P, P]),
    "rd_create_from_model": (C.c_int, [P, C.c_double, C.c_uint64, C.POINTER(P)]),
    "rd_set_composite": (C.c_int, [P, C.c_int, P, C.c_uint32]),
    "rd_set_composite_combine": (C.c_int, [P, C.c_int]),
    "rd_set_external_mode": (C.c_int, [P, C.c_int]),
    "rd_composite_eval": (C.c_int, [P, C.c_long, P, P, P]),
    "rd_surrogate_init": (C.c_int, [P, C.c_uint64]),
    "rd_surrogate_set_weights": (C.c_int, [P, P]),
    "rd_surrogate_get_weights": (C.c_int, [P, P]),
    "rd_surrogate_set_norm": (C.c_int, [P, P, P, C.c_double]),
    "rd_surrogate_get_norm": (C.c_int, [P, P, P, C.POINTER(C.c_double)]),
    "rd_surrogate_make_data": (C.c_int, [P, C.c_long, C.c_uint64, C.POINTER(C.c_double)]),
    "rd_surrogate_train": (C.c_int, [P, C.POINTER(rd_train_params), P]),
    "rd_surrogate_loss_grad": (C.c_int, [P, C.c_long, P, P, C.POINTER(C.c_double), P]),
    "rd_surrogate_predict": (C.c_int, [P, C.c_long, P, P]),
    "rd_surrogate_validate": (C.c_int, [P, C.c_long, C.c_uint64, C.POINTER(C.c_double)]),
    "rd_bench_steps": (C.c_int, [P, C.c_int, P]),
    "rd_bench_steps_direct": (C.c_int, [P, C.c_int, P]),
    "rd_bench_copy": (C.c_int, [P, C.c_long, C.c_int, C.POINTER(C.c_double)]),
    "rd_bench_ext": (C.c_int, [P, C.c_int, C.POINTER(C.c_double)]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list:
    return list(_SIGS) + ["Java_readdy_impl_sim_top_TopGPU_cCreateSimulation"]
