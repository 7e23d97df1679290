"""ctypes binding of ``libsovlp.so`` (declarations in ``include/sovlp.h``).

The library is built in-tree (``__graft_entry__.build()`` or ``make -C
sovereign-moe_amd/csrc``). Loading fails loudly when it is missing: there is no
CPU fallback anywhere in the package.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SOVLP_LIBRARY names another build of the same sources (e.g. one compiled
# with other -D options, for comparing them); the in-tree build otherwise
LIB_PATH = os.environ.get("SOVLP_LIBRARY") or os.path.join(HERE, "libsovlp.so")

_lib = None

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F64 = C.c_double


class SovlpError(RuntimeError):
    """A nonzero return code of the C ABI, with the library's message."""


class sovlp_params(C.Structure):
    _fields_ = [("eps_abs", F64), ("eps_rel", F64), ("feas_cap", F64), ("max_iter", I64),
                ("time_limit", F64), ("check_every", I32), ("stepsize_every", I32),
                ("beta_sufficient", F64), ("beta_necessary", F64), ("beta_artificial", F64),
                ("primal_weight_smoothing", F64), ("refine_rounds", I32), ("refine_stall", I32),
                ("refine_eps", F64), ("refine_tol", F64), ("refine_stall_open", I32),
                ("flags", I32)]


class sovlp_lp(C.Structure):
    _fields_ = [("m", I32), ("n", I32), ("nnz", I64),
                ("rp", P), ("ri", P), ("rx", P), ("cp", P), ("ci", P), ("cx", P),
                ("c", P), ("rl", P), ("ru", P), ("lo", P), ("hi", P), ("v0", P),
                ("tiles_a", P), ("ntiles_a", I32), ("tiles_t", P), ("ntiles_t", I32),
                ("obj_offset", F64), ("maximise", I32)]


class sovlp_info(C.Structure):
    _fields_ = [("status", I32), ("stage", I32), ("iterations", I64), ("restarts", I64),
                ("rejected_steps", I64), ("history_len", I64), ("primal_res", F64),
                ("dual_res", F64), ("gap", F64), ("primal_obj", F64), ("dual_obj", F64),
                ("step_size", F64), ("primal_weight", F64), ("objective", F64),
                ("scale_obj", F64), ("device_ticks", I64), ("refine_rows", I32),
                ("refine_round", I32), ("refine_scale", F64), ("start_tick", I64)]


class sovlp_certificate(C.Structure):
    _fields_ = [("bound", F64), ("perturbation", F64), ("objective", F64),
                ("col_violation", F64), ("row_violation", F64), ("vacuous", I32),
                ("reserved", I32)]


class sovlp_bnr_batch(C.Structure):
    _fields_ = [("m", I32), ("n", I32), ("K", I32), ("iters", I32), ("rp", P), ("cp", P),
                ("ri", P), ("ci", P), ("rx", P), ("cx", P), ("c", P), ("rl", P), ("ru", P),
                ("L", P), ("U", P), ("X", P), ("Y", P), ("eta", P), ("omega", P),
                ("step_growth", F64), ("step_damp", F64), ("bounds", P), ("empty", P),
                ("dual_sign", F64)]


class sovlp_host_lp(C.Structure):
    _fields_ = [("m", I32), ("n", I32), ("nnz", I64), ("cp", P), ("ci", P), ("cx", P),
                ("c", P), ("col_lb", P), ("col_ub", P), ("row_lb", P), ("row_ub", P)]


class sovlp_simplex_params(C.Structure):
    _fields_ = [("max_iter", I64), ("time_limit", F64), ("feas_tol", F64), ("opt_tol", F64),
                ("pivot_tol", F64), ("refactor_freq", I32), ("perturb_after", I32),
                ("perturb_scale", F64)]


class sovlp_simplex_result(C.Structure):
    _fields_ = [("status", I32), ("repairs", I32), ("iterations", I64), ("pivots", I64),
                ("phase1_iterations", I64), ("superbasics", I64), ("objective", F64),
                ("primal_infeasibility", F64),
                ("dual_infeasibility", F64), ("seconds", F64)]


# name -> (restype, argtypes); every entry point of include/sovlp.h
_SIGS = {
    "sovlp_last_error": (C.c_char_p, []),
    "sovlp_version": (C.c_int, []),
    "sovlp_selftest": (C.c_int, []),
    "sovlp_clock_rate": (C.c_int, [C.c_int, C.POINTER(F64)]),
    "sovlp_spmv_csr_f64": (C.c_int, [C.c_int, P, P, P, P, P, C.c_int, P]),
    "sovlp_spmv_csr_f32": (C.c_int, [C.c_int, P, P, P, P, P, C.c_int, P]),
    "sovlp_primal_step_f64": (C.c_int, [P, P, P, P, P, P, F64, C.c_int, P]),
    "sovlp_dual_step_f64": (C.c_int, [P, P, P, P, P, F64, C.c_int, P]),
    "sovlp_pdhg_primal_f64": (C.c_int, [P, P, P, P, P, P, F64, C.c_int, P]),
    "sovlp_pdhg_dual_f64": (C.c_int, [P, P, P, P, P, P, F64, C.c_int, P]),
    "sovlp_pdhg_accum_f64": (C.c_int, [P, P, P, P, F64, C.c_int, P]),
    "sovlp_dot_f64": (C.c_int, [P, P, I64, P, P]),
    "sovlp_nrm2_f64": (C.c_int, [P, I64, P, P]),
    "sovlp_nrminf_f64": (C.c_int, [P, I64, P, P]),
    "sovlp_dot_f32": (C.c_int, [P, P, I64, P, P]),
    "sovlp_nrm2_f32": (C.c_int, [P, I64, P, P]),
    "sovlp_plan_tiles": (C.c_int, [C.c_int, P, C.c_int, C.c_int, C.c_int, P, C.POINTER(I32)]),
    "sovlp_partition_tiles": (C.c_int, [C.c_int, P, P, C.c_int, P]),
    "sovlp_tile_geometry": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "sovlp_spmv_tiled_f64": (C.c_int, [C.c_int, P, P, P, P, C.c_int, P, C.c_int, P, P, P]),
    "sovlp_spmv_tiled_f32": (C.c_int, [C.c_int, P, P, P, P, C.c_int, P, C.c_int, P, P, P]),
    "sovlp_spmv_panels_f64": (C.c_int, [C.c_int, C.c_int, P, P, P, P, P, P, C.c_int, P, P, P,
                                        P]),
    "sovlp_spmv_panels_f32": (C.c_int, [C.c_int, C.c_int, P, P, P, P, P, P, C.c_int, P, P, P,
                                        P]),
    "sovlp_pdlp_workspace_bytes": (C.c_int, [C.POINTER(sovlp_lp), C.c_int, I64, C.POINTER(I64)]),
    "sovlp_pdlp_create": (C.c_int, [C.POINTER(sovlp_lp), C.POINTER(sovlp_params), C.c_int, P,
                                    I64, I64, C.c_int, P, C.POINTER(P)]),
    "sovlp_pdlp_advance": (C.c_int, [P, I64, F64, P]),
    "sovlp_pdlp_info": (C.c_int, [P, C.POINTER(sovlp_info), P]),
    "sovlp_pdlp_outputs": (C.c_int, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "sovlp_pdlp_width": (C.c_int, [P, C.POINTER(C.c_int)]),
    "sovlp_pdlp_phase_ticks": (C.c_int, [P, P, P]),
    "sovlp_batch_phase_ticks": (C.c_int, [P, P, P]),
    "sovlp_pdlp_scaling": (C.c_int, [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "sovlp_phases_per_launch": (C.c_int, [C.POINTER(C.c_int)]),
    "sovlp_pdlp_refinement": (C.c_int, [P, C.POINTER(P)]),
    "sovlp_pdlp_destroy": (C.c_int, [P]),
    "sovlp_phase_slots": (C.c_int, [C.POINTER(C.c_int)]),
    "sovlp_phase_name": (C.c_char_p, [C.c_int]),
    "sovlp_pdlp_phase_counts": (C.c_int, [P, P, P]),
    "sovlp_batch_geometry": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sovlp_lds_resident": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, I64, C.c_int, C.c_int,
                                     C.POINTER(C.c_int)]),
    "sovlp_batch_workspace_bytes": (C.c_int, [C.c_int, P, C.c_int, I64, C.POINTER(I64)]),
    "sovlp_batch_create": (C.c_int, [C.c_int, P, C.POINTER(sovlp_params), C.c_int, C.c_int, P,
                                     I64, I64, P, C.POINTER(P)]),
    "sovlp_batch_create_mixed": (C.c_int, [C.c_int, P, C.POINTER(sovlp_params), C.c_int, P, P,
                                           I64, I64, P, C.POINTER(P)]),
    "sovlp_batch_run": (C.c_int, [P, F64, I64, P, C.POINTER(I64)]),
    "sovlp_batch_infos": (C.c_int, [P, P, P]),
    "sovlp_batch_outputs": (C.c_int, [P, C.c_int, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "sovlp_batch_refinement": (C.c_int, [P, C.c_int, C.POINTER(P)]),
    "sovlp_batch_phase_counts": (C.c_int, [P, P, P]),
    "sovlp_batch_destroy": (C.c_int, [P]),
    "sovlp_server_workspace_bytes": (C.c_int, [C.c_int, C.c_int, C.POINTER(I64)]),
    "sovlp_server_team_workspace_bytes": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(I64)]),
    "sovlp_server_lp_bytes": (C.c_int, [C.POINTER(sovlp_lp), C.c_int, I64, C.POINTER(I64)]),
    "sovlp_server_create_team": (C.c_int, [C.c_int, C.c_int, C.POINTER(sovlp_params), C.c_int,
                                           C.c_int, P, I64, P, C.POINTER(P)]),
    "sovlp_server_create": (C.c_int, [C.c_int, C.POINTER(sovlp_params), C.c_int, C.c_int,
                                      C.c_int, P, I64, P, C.POINTER(P)]),
    "sovlp_server_admit": (C.c_int, [P, C.c_int, C.POINTER(sovlp_lp), P, P, F64, P, I64, I64,
                                     P]),
    "sovlp_server_step": (C.c_int, [P, C.c_int, P, P, C.POINTER(I32)]),
    "sovlp_server_info": (C.c_int, [P, C.c_int, C.POINTER(sovlp_info)]),
    "sovlp_server_outputs": (C.c_int, [P, C.c_int, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "sovlp_server_refinement": (C.c_int, [P, C.c_int, C.POINTER(P)]),
    "sovlp_server_history": (C.c_int, [P, C.c_int, C.POINTER(P)]),
    "sovlp_server_phase_counts": (C.c_int, [P, C.c_int, P, P]),
    "sovlp_server_release": (C.c_int, [P, C.c_int]),
    "sovlp_server_clock_hz": (C.c_int, [P, C.POINTER(F64)]),
    "sovlp_server_destroy": (C.c_int, [P]),
    "sovlp_residual_workspace_bytes": (C.c_int, [C.c_int, C.POINTER(I64)]),
    "sovlp_bnr_workspace_bytes": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(I64)]),
    "sovlp_bnr_solve_batch": (C.c_int, [C.POINTER(sovlp_bnr_batch), P, I64, P]),
    "sovlp_spmm_csr_f64": (C.c_int, [C.c_int, P, P, P, P, P, C.c_int, P]),
    "sovlp_spmm_csr_f32": (C.c_int, [C.c_int, P, P, P, P, P, C.c_int, P]),
    "sovlp_residual_f64": (C.c_int, [C.c_int, P, P, P, P, P, P, P, I64, P]),
    "sovlp_residual_noalloc_max_rows": (C.c_int, []),
    "sovlp_residual_f64_noalloc": (C.c_int, [C.c_int, P, P, P, P, P, P, P]),
    "sovlp_certify_workspace_bytes": (C.c_int, [C.c_int, C.c_int, C.POINTER(I64)]),
    "sovlp_certify_f64": (C.c_int, [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P, P, F64,
                                    C.c_int, P, P, F64, P, I64, P, P,
                                    C.POINTER(sovlp_certificate), P]),
    "sovlp_simplex_solve": (C.c_int, [C.POINTER(sovlp_host_lp),
                                      C.POINTER(sovlp_simplex_params), P, P, P, P, P,
                                      C.POINTER(sovlp_simplex_result)]),
    "sovlp_basis_from_point": (C.c_int, [C.POINTER(sovlp_host_lp), P, P, F64, F64, P,
                                         C.POINTER(I32)]),
    "sovlp_basis_repairs": (C.c_int, [C.POINTER(sovlp_host_lp), P, C.POINTER(I32)]),
    "sovlp_simplex_ranging": (C.c_int, [C.POINTER(sovlp_host_lp), P, P, P, P, P, P, P, P, P,
                                        P]),
    "sovlp_crossover": (C.c_int, [C.POINTER(sovlp_host_lp), C.POINTER(sovlp_simplex_params),
                                  P, P, P, P, P, P, C.POINTER(sovlp_simplex_result)]),
}


def exported_symbols():
    return sorted(_SIGS)


def lib():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsovlp.so is not built ({LIB_PATH}); run "
                               "`python -c 'import __graft_entry__ as g; g.build()'` "
                               "or `make -C sovereign-moe_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().sovlp_last_error().decode(errors="replace")
        raise SovlpError(f"{what or 'sovlp'} failed (code {rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
