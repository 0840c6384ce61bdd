"""ctypes binding of libhfx_hip.so (C ABI in include/hfx_hip.h)."""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libhfx_hip.so')


class HfxError(RuntimeError):
    """An hfx_* entry point returned a nonzero status."""


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError('libhfx_hip.so is not built (%s); run __graft_entry__.build()' % LIB_PATH)
    return ctypes.CDLL(LIB_PATH)


lib = _load()

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong

_SIGS = {
    'hfx_last_error': ([], ctypes.c_char_p),
    'hfx_device_count': ([_P], _I),
    'hfx_refel_sizes': ([_I, _I, _P], _I),
    'hfx_refel_tables': ([_I, _I] + [_P] * 9, _I),
    'hfx_cubature': ([_I, _I, _P, _P], _I),
    'hfx_mesh_structured': ([_I, _I, _I, _D, ctypes.c_ulonglong, _P], _I),
    'hfx_mesh_from_cells': ([_I, _I, _I, _P, _I, _P, _P], _I),
    'hfx_mesh_structured_gpu': ([_I, _I, _I, _I, _D, ctypes.c_ulonglong, _P], _I),
    'hfx_mesh_from_cells_gpu': ([_I, _I, _I, _I, _P, _I, _P, _P], _I),
    'hfx_mesh_sizes': ([_P, _P], _I),
    'hfx_mesh_get': ([_P] * 8, _I),
    'hfx_mesh_free': ([_P], None),
    'hfx_hdg_create': ([_I, _I, _I, _P], _I),
    'hfx_hdg_destroy': ([_P], None),
    'hfx_hdg_set_mesh': ([_P, _P], _I),
    'hfx_hdg_set_model': ([_P, _I], _I),
    'hfx_hdg_set_fields': ([_P, _P, _I, _P, _I, _P, _P], _I),
    'hfx_hdg_ip_coords': ([_P, _P], _I),
    'hfx_hdg_set_source': ([_P, _P, _P], _I),
    'hfx_hdg_set_launch': ([_P, _I, _I], _I),
    'hfx_hdg_assemble': ([_P], _I),
    'hfx_hdg_local_matrices': ([_P, _I, _I, _P, _P], _I),
    'hfx_hdg_get_condensed': ([_P] * 7, _I),
    'hfx_hdg_trace_sizes': ([_P, _P], _I),
    'hfx_hdg_get_trace_csr': ([_P] * 5, _I),
    'hfx_hdg_solve': ([_P, _I, _D, _I, _I, _P, _P], _I),
    'hfx_hdg_solve_noalloc': ([_P, _I, _D, _I, _P, _P], _I),
    'hfx_hdg_pc_apply': ([_P, _I, _P, _P], _I),
    'hfx_hdg_set_fused_scatter': ([_P, _I], _I),
    'hfx_hdg_set_time_scheme': ([_P, _I, _D], _I),
    'hfx_hdg_set_solution': ([_P, _P], _I),
    'hfx_hdg_set_solver_type': ([_P, _I], _I),
    'hfx_hdg_set_flux': ([_P, _P], _I),
    'hfx_hdg_set_butcher': ([_P, _I, _P, _I], _I),
    'hfx_rk_table': ([_I, _P, _P], _I),
    'hfx_hdg_rk_start_step': ([_P], _I),
    'hfx_hdg_rk_compute_stage': ([_P], _I),
    'hfx_hdg_rk_compute_solution': ([_P], _I),
    'hfx_hdg_rk_stages': ([_P, _P], _I),
    'hfx_hdg_set_trace': ([_P, _P], _I),
    'hfx_hdg_set_boundary_model': ([_P, _I, _P, _I], _I),
    'hfx_hdg_set_boundary_hdg': ([_P, _P, _I, _P, _P, _I, _I], _I),
    'hfx_hdg_set_solver_graphs': ([_P, _I], _I),
    'hfx_hdg_ilu_factors': ([_P, _I, _P, _P, _P], _I),
    'hfx_ilu_analyse': ([_I, _P, _P, _I, _P, _P, _P, _P, _P], _I),
    'hfx_hdg_get_solution': ([_P] * 4, _I),
    'hfx_hdg_time': ([_P, _I, _I, _I, _P], _I),
    'hfx_element_lds_bytes': ([_I, _I, _P], _I),
    'hfx_element_workspace_bytes': ([_I, _I, _P], _I),
    'hfx_hdg_set_kernel': ([_P, _I], _I),
    'hfx_hdg_set_krylov': ([_P, _I], _I),
    'hfx_hdg_kernel_info': ([_P, _P], _I),
    'hfx_la_create': ([_I, _P], _I),
    'hfx_la_free': ([_P], None),
    'hfx_la_initialize': ([_P], _I),
    'hfx_la_configure': ([_P, _I, _I, _D, _I, _I], _I),
    'hfx_la_allocate': ([_P, _I, _P, _P], _I),
    'hfx_la_add_vals': ([_P, _I, _P, _I, _P, _P], _I),
    'hfx_la_set_vals': ([_P, _I, _P, _I, _P, _P], _I),
    'hfx_la_add_rhs': ([_P, _I, _P, _P], _I),
    'hfx_la_set_rhs': ([_P, _I, _P, _P], _I),
    'hfx_la_zero_rows': ([_P, _I, _P], _I),
    'hfx_la_assemble': ([_P], _I),
    'hfx_la_assemble_flush': ([_P], _I),
    'hfx_la_solve': ([_P, _P, _I, _P, _P], _I),
    'hfx_la_ownership': ([_P, _P, _P], _I),
    'hfx_la_clear': ([_P], _I),
    'hfx_la_destroy_system': ([_P], _I),
    'hfx_la_num_dofs': ([_P, _P], _I),
    'hfx_la_get_values': ([_P, _I, _P, _I, _P, _P], _I),
    'hfx_la_get_rhs': ([_P, _I, _P, _P], _I),
    'hfx_la_nnz': ([_P, _P], _I),
    'hfx_la_get_csr': ([_P] * 5, _I),
    'hfx_batch_create': ([_I, _P], _I),
    'hfx_batch_free': ([_P], None),
    'hfx_batch_add': ([_P, _P], _I),
    'hfx_batch_size': ([_P, _P, _P], _I),
    'hfx_batch_set_config': ([_P, _I, _I, _I], _I),
    'hfx_batch_get_config': ([_P, _P], _I),
    'hfx_batch_tune': ([_P, _I, _P], _I),
    'hfx_batch_assemble': ([_P, _I], _I),
    'hfx_batch_time': ([_P, _I, _I, _I, _P], _I),
    'hfx_lpc_create': ([_I, _I, _P], _I),
    'hfx_lpc_free': ([_P], None),
    'hfx_lpc_num_params': ([_P, _P], _I),
    'hfx_lpc_init_params': ([_P, ctypes.c_ulonglong], _I),
    'hfx_lpc_set_params': ([_P, _P, ctypes.c_longlong], _I),
    'hfx_lpc_get_params': ([_P, _P, ctypes.c_longlong], _I),
    'hfx_lpc_add_sample': ([_P, _P], _I),
    'hfx_lpc_num_samples': ([_P, _P], _I),
    'hfx_lpc_clear_samples': ([_P], _I),
    'hfx_lpc_graph_sizes': ([_P, _I, _P], _I),
    'hfx_lpc_loss_grad': ([_P, _I, _I, ctypes.c_ulonglong, _P, _P], _I),
    'hfx_lpc_train': ([_P, _P, _I, ctypes.c_float, _I, ctypes.c_ulonglong, _P], _I),
    'hfx_lpc_outputs': ([_P, _I, _I, _P, _P], _I),
    'hfx_hdg_set_lpc': ([_P, _P, _I], _I),
    'hfx_hdg_pc_setup_ms': ([_P, _P], _I),
    'hfx_burgers_create': ([_I, _I, _I, _P], _I),
    'hfx_burgers_destroy': ([_P], None),
    'hfx_burgers_set_mesh': ([_P, _P], _I),
    'hfx_burgers_set_fields': ([_P, _P, _I, _P, _P, _P], _I),
    'hfx_burgers_set_boundary_model': ([_P, _I], _I),
    'hfx_ngamma_create': ([_I, _I, _P], _I),
    'hfx_multi_create': ([_I, _I, _I, _P], _I),
    'hfx_burgers_set_boundary_hdg': ([_P, _P, _I, _P, _P, _I, _I], _I),
    'hfx_ngamma_set_boundary_bohm': ([_P, _P, _I, _P, _I], _I),
    'hfx_burgers_boundary_rows': ([_P, _P, _P], _I),
    'hfx_burgers_set_time_scheme': ([_P, _I, _D], _I),
    'hfx_burgers_set_solver_graphs': ([_P, _I], _I),
    'hfx_burgers_start_step': ([_P], _I),
    'hfx_burgers_set_dirichlet': ([_P, _P], _I),
    'hfx_ngamma_set_fields': ([_P, _P, _I, _P, _P, _P, _D, _P, _P], _I),
    'hfx_burgers_set_krylov': ([_P, _I], _I),
    'hfx_burgers_set_state': ([_P, _P, _P, _P], _I),
    'hfx_burgers_get_state': ([_P, _P, _P, _P, _P], _I),
    'hfx_burgers_sizes': ([_P, _P], _I),
    'hfx_burgers_assemble': ([_P], _I),
    'hfx_burgers_condensed': ([_P] * 7, _I),
    'hfx_burgers_trace_system': ([_P] * 5, _I),
    'hfx_burgers_solve': ([_P, _I, _D, _I, _I, _P, _P, _P], _I),
    'hfx_burgers_nonlinear_solve': ([_P, _I, _D, _D, _I, _D, _I, _I, _P, _P, _P], _I),
}

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res


def check(status):
    """Raise HfxError with hfx_last_error() if status is nonzero."""
    if status != 0:
        raise HfxError(lib.hfx_last_error().decode())
    return status
