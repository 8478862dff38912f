"""ctypes binding of libtinymd_amd.so (include/md_abi.h + include/tmd_ext.h).

The library is the product: it must be built (``make -C molecular-dynamics_amd``
or ``__graft_entry__.build()``).  Loading fails loudly when it is missing;
there is no CPU fallback.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(ROOT, "lib", "libtinymd_amd.so")


class Vector3D(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double)]


_d = ctypes.c_double
_i = ctypes.c_int
_b = ctypes.c_bool
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.POINTER(Vector3D)
_llp = ctypes.POINTER(ctypes.c_longlong)
_ulp = ctypes.POINTER(ctypes.c_ulong)

# name -> (restype, argtypes); every entry point of include/md_abi.h and
# include/tmd_ext.h
SIGNATURES = {
    # md_abi.h (anydsl_includes.h:10-48)
    "md_initialize_grid": (_i, [_dp, _vp, _vp, _i, _dp, _dp, _d, _i, _i]),
    "md_rescale_grid": (None, [_dp]),
    "md_copy_data_from_accelerator": (None, []),
    "md_copy_data_to_accelerator": (None, []),
    "md_initial_integration": (None, [_d]),
    "md_final_integration": (None, [_d]),
    "md_enforce_pbc": (None, []),
    "md_write_grid_data_to_arrays": (_i, [_dp, _vp, _vp, _vp]),
    "md_write_grid_ghost_data_to_arrays": (_i, [_dp, _vp, _vp, _vp]),
    "md_write_grid_aabb_data_to_arrays": (_i, [_dp, _vp, _vp, _vp]),
    "md_distribute_particles": (None, []),
    "md_assemble_neighborlists": (None, [_b, _d]),
    "md_deallocate_grid": (None, []),
    "md_print_grid": (None, []),
    "md_print_ghost": (None, []),
    "md_create_velocity": (None, [_d]),
    "md_compute_lennard_jones": (None, [_b, _d, _d, _d]),
    "md_compute_dem": (None, [_b, _d, _d, _d, _d, _d]),
    "md_mpi_initialize": (None, []),
    "md_mpi_finalize": (None, []),
    "md_synchronize_ghost_layer": (None, []),
    "md_borders": (None, []),
    "md_exchange_particles": (None, []),
    "md_get_send_buffer_value": (_d, [_i]),
    "md_set_recv_buffer_value": (None, [_i, _d]),
    "md_resize_recv_buffer_capacity": (None, [_i]),
    "md_serialize_particles": (_i, [_dp]),
    "md_deserialize_particles": (None, [_i]),
    "md_get_world_size": (_i, []),
    "md_get_world_rank": (_i, []),
    "md_get_number_of_particles": (_i, []),
    "md_get_number_of_ghost_particles": (_i, []),
    "md_get_node_bounding_box": (None, [_dp, _dp]),
    "md_compute_boundary_weights": (None, [_d, _d, _d, _d, _d, _d, _ulp, _ulp]),
    "md_update_neighborhood": (None, [_i, _i, _ip, _ip, _dp]),
    "md_report_iterations": (None, []),
    "md_report_particles": (None, []),
    "md_report_time": (None, [_d, _d, _d, _d, _d, _d]),
    "md_barrier": (None, []),
    # tmd_ext.h
    "tmd_last_error": (ctypes.c_char_p, []),
    "tmd_clear_error": (None, []),
    "tmd_device_info": (_i, [ctypes.c_char_p, _i]),
    "tmd_kahan_sum": (_d, [ctypes.POINTER(ctypes.c_float), _i]),
    "tmd_set_precision": (None, [_i]),
    "tmd_set_adjoint_precision": (None, [_i]),
    "tmd_set_sort_particles": (None, [_i]),
    "tmd_set_half_mode": (None, [_i]),
    "tmd_set_ghost_mode": (None, [_i]),
    "tmd_set_sync": (None, [_i]),
    "tmd_set_dem_radius": (None, [_d]),
    "tmd_set_lj_shift": (None, [_i]),
    "tmd_set_tile": (None, [_i, _i, _i]),
    "tmd_set_tile_threads": (None, [_i]),
    "tmd_thermo": (None, [_d, _d, _d, _dp]),
    "tmd_thermo_eshift": (None, [_d, _d, _d, _dp]),
    "tmd_write_vtk": (_i, [ctypes.c_char_p, _i]),
    "tmd_save_state": (_i, [ctypes.c_char_p]),
    "tmd_load_state": (_i, [ctypes.c_char_p]),
    "tmd_thermo_noalloc": (None, [_d, _d, _d, _dp]),
    "tmd_get_grid_info": (None, [_ip]),
    "tmd_get_tile_info": (None, [_ip]),
    "tmd_get_ghost_table": (_i, [_ip, _ip, _ip, _ip]),
    "tmd_get_cells": (_i, [_ip, _ip, _ip]),
    "tmd_get_neighbor_lists": (_i, [_ip, _ip, _i]),
    "tmd_run_lj": (_i, [_i, _i, _d, _b, _d, _d, _d, _d]),
    "tmd_run_dem": (_i, [_i, _i, _d, _b, _d, _d, _d, _d, _d, _d]),
    "tmd_profile_enable": (None, [_i]),
    "tmd_profile_get": (None, [_i, _dp, _llp]),
    "tmd_profile_reset": (None, []),
    "tmd_generate_fcc": (_i, [_i, _i, _i, _d, _i, _dp, _vp, _vp, _i]),
    "tmd_generate_granular_gas": (_i, [_i, _d, ctypes.c_ulonglong, _d, _dp, _vp, _vp, _i]),
    "tmd_generate_replica_set": (_i, [ctypes.c_ulonglong, _i, _i, _i, _d, _d, _d, _d, _ip, _dp, _dp]),
    "tmd_batch_create": (_i, [_i, _ip, _vp, _vp, _dp]),
    "tmd_batch_setup": (_i, [_i, _d, _d, _d, _d, _d, _i, _i, _i]),
    "tmd_batch_configure": (_i, [_i, _i, _i, _d, _d, _d]),
    "tmd_batch_autotune": (_i, [_i, _i, _dp]),
    "tmd_batch_reset": (_i, [_i]),
    "tmd_batch_set_graph": (_i, [_i, _i]),
    "tmd_batch_run": (_i, [_i, _i]),
    "tmd_batch_run_timed": (_i, [_i, _i, _dp]),
    "tmd_batch_get_state": (_i, [_i, _i, _vp, _vp, _vp]),
    "tmd_batch_thermo": (_i, [_i, _dp]),
    "tmd_batch_info": (_i, [_i, _ip]),
    "tmd_batch_destroy": (None, [_i]),
    "tmd_traj_loss": (_i, [_i, _i, _d, _d, _d, _d, _d, _dp]),
    "tmd_traj_adjoint": (_i, [_i, _i, _d, _d, _d, _d, _d, _d, _d, _dp, _dp, _vp, _vp]),
}

_lib = None


class TinyMDError(RuntimeError):
    pass


def load():
    """Load the HIP library (built in-tree).  Raises when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libtinymd_amd.so not built ({LIB_PATH}); run `make -C molecular-dynamics_amd` "
                "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check():
    """Raise TinyMDError if the library recorded an error."""
    err = load().tmd_last_error()
    if err:
        raise TinyMDError(err.decode())
