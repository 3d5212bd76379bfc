"""ctypes declarations of include/*.h (engine, builder, batched sweep), and the loader.

This is the binding a Python caller (or a test playing the reference's Rust caller) uses to
reach libfieldgrid.so. Loading fails loudly when the library has not been built: there is
no CPU fallback behind it.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfieldgrid.so")

FG_OK, FG_CANCELLED, FG_UNSTABLE, FG_TOO_LARGE, FG_UNSUPPORTED, FG_HIP, FG_INVALID = range(7)
FG_RESISTOR, FG_CAPACITOR, FG_INDUCTOR = 0, 1, 2
FG_ENGINE_AUTO, FG_ENGINE_SPLIT, FG_ENGINE_FUSED = 0, 1, 2
FG_MAX_FREQS = 4

u32, u64, i64, f32, f64 = C.c_uint32, C.c_uint64, C.c_int64, C.c_float, C.c_double
P = C.POINTER


class fg_port(C.Structure):
    _fields_ = [("comp", u32), ("first_edge", u32), ("n_edges", u32), ("first_loop", u32),
                ("n_loops", u32), ("first_source", u32), ("n_sources", u32), ("pad", u32)]


class fg_port_edge(C.Structure):
    _fields_ = [("id", u64), ("w", f32), ("pad", u32)]


class fg_loop_term(C.Structure):
    _fields_ = [("id", u64), ("comp", u32), ("w", f32)]


class fg_source(C.Structure):
    _fields_ = [("id", u64), ("comp", u32), ("src", f32)]


class fg_inductor(C.Structure):
    _fields_ = [("id", u64), ("comp", u32), ("k_e", f32), ("k_i", f32), ("i0", f32)]


class fg_debye_edge(C.Structure):
    _fields_ = [("f", u64), ("dd", f32), ("set", u32)]


class fg_wire_h(C.Structure):
    _fields_ = [("h", u64), ("e", u64 * 4), ("c", f32 * 4)]


class fg_problem(C.Structure):
    _fields_ = [
        ("n", u32 * 3), ("pml", u32), ("dt", f64), ("axes", P(f32)),
        ("coef_wide", u32), ("n_sets", u32), ("coef_sets", P(f32)), ("coef_index", P(u32)),
        ("n_ports", u32), ("n_sources", u32), ("n_port_edges", u32), ("n_loops", u32),
        ("ports", P(fg_port)), ("port_edges", P(fg_port_edge)), ("loops", P(fg_loop_term)),
        ("sources", P(fg_source)),
        ("n_inductors", u32), ("pad0", u32), ("inductors", P(fg_inductor)),
        ("n_debye", u64), ("debye_edges", P(fg_debye_edge)), ("n_poles", u32), ("n_pole_sets", u32),
        ("debye_table", P(f32)),
        ("n_sheets", u64), ("sheets", P(f32)), ("sheet_states", u32), ("n_branches", u32),
        ("branches", P(f32)), ("n_wires", u64), ("wires", P(fg_wire_h)),
    ]


class fg_pulse(C.Structure):
    _fields_ = [("f0", f64), ("fc", f64)]


class fg_point_probe(C.Structure):
    _fields_ = [("field", u32), ("i", u32), ("j", u32), ("k", u32)]


class fg_extras(C.Structure):
    _fields_ = [
        ("min_steps", u64), ("max_steps", u64), ("decay_db", f64), ("nf", u32), ("engine", u32),
        ("freqs", f64 * FG_MAX_FREQS), ("plane_k", i64), ("n_patches", u64), ("ntff", P(u32)),
        ("check_every", u64), ("profile", u32), ("n_probes", u32), ("probes", P(fg_point_probe)),
    ]


class fg_record(C.Structure):
    _fields_ = [
        ("steps", u64), ("ports", u32), ("pad", u32), ("decay_db", f64),
        ("series", P(f32)), ("series_cap", u64), ("plane", P(f32)), ("plane_cap", u64),
        ("ntff", P(f32)), ("ntff_cap", u64), ("probes", P(f32)), ("probes_cap", u64),
        ("e_out", P(f32)), ("h_out", P(f32)),
    ]


class fg_edge(C.Structure):
    _fields_ = [("comp", u32), ("at", u32 * 3)]


class fg_lumped(C.Structure):
    _fields_ = [("kind", u32), ("n_edges", u32), ("value", f64), ("edges", P(fg_edge))]


class fg_port_def(C.Structure):
    _fields_ = [("r", f64), ("reference_above", u32), ("n_columns", u32), ("column_len", P(u32)),
                ("edges", P(fg_edge))]


class fg_resistive(C.Structure):
    _fields_ = [("comp", u32), ("adaptive", u32), ("id", u64), ("r", f64), ("faces", u64 * 2)]


class fg_wire(C.Structure):
    _fields_ = [("edge", fg_edge), ("radius", f64)]


class fg_sim_desc(C.Structure):
    _fields_ = [
        ("line", P(f64) * 3), ("len", u32 * 3), ("pml", u32),
        ("eps", P(f32)), ("sigma", P(f32)), ("inf", P(f32)), ("debye", P(f32)), ("pec", P(C.c_uint8)),
        ("n_lumped", u32), ("n_ports", u32), ("lumped", P(fg_lumped)), ("ports", P(fg_port_def)),
        ("n_resistive", u32), ("n_wires", u32), ("resistive", P(fg_resistive)), ("wires", P(fg_wire)),
        ("n_poles", u32), ("n_pole_sets", u32), ("poles", P(f64)), ("set_first", P(u32)),
        ("debye_set", P(C.c_uint8)),
        ("sheet_states", u32), ("sheet_table_len", u32), ("sheet_table", P(f32)),
    ]


class fg_tune(C.Structure):
    _fields_ = [("tj", u32), ("seg", u32), ("chunk", u32), ("depth", u32), ("cache", C.c_char_p),
                ("used_tj", u32), ("used_seg", u32), ("used_chunk", u32), ("from_cache", u32),
                ("used_depth", u32), ("kept_steps", u32), ("tune_ms", f64), ("step_ms", f64),
                ("launch", u32), ("used_launch", u32)]


class fg_spectra(C.Structure):
    _fields_ = [("nf", u32), ("pad", u32), ("freqs", P(f64)), ("v", P(f64)), ("i", P(f64))]


class fg_design(C.Structure):
    _fields_ = [("lo", u32 * 3), ("hi", u32 * 3), ("n_obs", u32), ("every", u32), ("obs", P(fg_point_probe)),
                ("steps", u64), ("whole_tape", u32), ("pad", u32)]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, u64, f64)

# every symbol the headers declare, with its ctypes signature
SIGNATURES = {
    "fg_open": (C.c_int, [C.c_int, P(C.c_void_p)]),
    "fg_close": (None, [C.c_void_p]),
    "fg_name": (C.c_char_p, [C.c_void_p]),
    "fg_max_cells": (u64, [C.c_void_p]),
    "fg_run": (C.c_int, [C.c_void_p, P(fg_problem), u32, P(fg_pulse), P(fg_extras), PROGRESS_FN,
                         C.c_void_p, P(fg_record)]),
    "fg_last_error": (C.c_char_p, []),
    "fg_session_new": (C.c_int, [C.c_void_p, P(fg_problem), u32, P(fg_pulse), P(fg_extras),
                                 P(C.c_void_p)]),
    "fg_session_step": (C.c_int, [C.c_void_p, u64]),
    "fg_session_sync": (C.c_int, [C.c_void_p]),
    "fg_session_timing": (C.c_int, [C.c_void_p, C.c_int, P(f64), P(u64)]),
    "fg_session_energy": (C.c_int, [C.c_void_p, P(f64)]),
    "fg_session_read": (C.c_int, [C.c_void_p, P(fg_record)]),
    "fg_session_free": (None, [C.c_void_p]),
    "fg_device_synchronize": (C.c_int, [C.c_void_p]),
    "fg_stream_copy": (C.c_int, [C.c_void_p, u64, u32, P(f64)]),
    "fg_sim_new": (C.c_int, [P(fg_sim_desc), P(C.c_void_p)]),
    "fg_sim_free": (None, [C.c_void_p]),
    "fg_sim_dt": (f64, [C.c_void_p]),
    "fg_sim_dims": (None, [C.c_void_p, P(u32)]),
    "fg_sim_ca": (P(f32), [C.c_void_p]),
    "fg_sim_cb": (P(f32), [C.c_void_p]),
    "fg_sim_counts": (None, [C.c_void_p, P(u64)]),
    "fg_sim_pack": (C.c_int, [C.c_void_p, u64, P(fg_problem)]),
    "fg_patches": (u64, [C.c_void_p, u32, P(u32), P(f64), P(f64)]),
    "fg_run_batch": (C.c_int, [C.c_void_p, u32, P(P(fg_problem)), P(u32), P(fg_pulse), P(P(fg_extras)),
                               P(fg_tune), PROGRESS_FN, C.c_void_p, P(fg_record), P(fg_spectra)]),
    "fg_adjoint_new": (C.c_int, [C.c_void_p, P(fg_sim_desc), P(fg_design), u32, P(fg_pulse), P(C.c_void_p)]),
    "fg_adjoint_forward": (C.c_int, [C.c_void_p, P(f32), P(f32)]),
    "fg_adjoint_backward": (C.c_int, [C.c_void_p, P(f64), P(f64)]),
    "fg_adjoint_forward64": (C.c_int, [C.c_void_p, P(f64), P(f64)]),
    "fg_adjoint_timing": (C.c_int, [C.c_void_p, P(f64), P(u64)]),
    "fg_adjoint_free": (None, [C.c_void_p]),
    "fg_session_steps_per_pass": (C.c_int, [C.c_void_p, P(u32)]),
    "fg_port_spectra": (C.c_int, [C.c_void_p, P(f32), u64, u32, f64, u32, P(f64), P(f64), P(f64)]),
    "fg_problem_check": (C.c_int, [P(fg_problem), u32, P(fg_extras)]),
    "fg_port_spectra_static": (C.c_int, [C.c_void_p, P(f32), u64, u32, f64, u32, P(f64), P(f64), P(f64)]),
    "fg_far_new": (C.c_int, [C.c_void_p, u64, P(f64), P(f64), P(f32), u32, u32, f64, f64, f64, P(C.c_void_p)]),
    "fg_far_field": (C.c_int, [C.c_void_p, u64, P(f64), P(f64)]),
    "fg_far_intensity": (C.c_int, [C.c_void_p, u64, P(f64), P(f64)]),
    "fg_far_radiated": (C.c_int, [C.c_void_p, u32, P(f64), P(f64)]),
    "fg_far_free": (None, [C.c_void_p]),
}

_lib = None


class FieldgridError(RuntimeError):
    """A non-zero fg_status, with the library's message (the reference's Err(String))."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def lib():
    """Load libfieldgrid.so once. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m fieldgrid_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != FG_OK:
        raise FieldgridError(status, lib().fg_last_error().decode())
    return status
