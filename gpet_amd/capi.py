"""ctypes mirror of include/gpet.h.

Struct layouts here must match the C header byte for byte; tests check the
sizes against the library.  Event and trace records are also exposed as
numpy dtypes so result arrays can be handled without copies.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

GPET_MAXMAT = 10
GPET_NTAB = 4096
GPET_NWCK = 4096
GPET_NCP = 301
GPET_NE = 151
GPET_MAXSURFACE = 8
GPET_MAXPANELS = 256
GPET_PATH_MAX = 256

STATUS = {0: "ok", -1: "bad argument", -2: "file missing or unreadable", -3: "format or dimension mismatch",
          -4: "device error", -5: "capacity exceeded", -6: "call order", -7: "unsupported"}

EVENT_DTYPE = np.dtype([("parn", "<i4"), ("pann", "<i4"), ("modn", "<i4"), ("cryn", "<i4"), ("siten", "<i4"),
                        ("eventid", "<i4"), ("t", "<f8"), ("E", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
assert EVENT_DTYPE.itemsize == 48

TRACE_DTYPE = np.dtype([("history", "<i8"), ("photon", "<i4"), ("status", "<i4"), ("n_steps", "<i4"),
                        ("n_compton", "<i4"), ("n_rayleigh", "<i4"), ("panel", "<i4"), ("n_steps_panel", "<i4"),
                        ("n_int_panel", "<i4"), ("n_events", "<i4"), ("draws", "<u4"),
                        ("x0", "<f4"), ("y0", "<f4"), ("z0", "<f4"), ("E0", "<f4"),
                        ("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4"),
                        ("E", "<f4"), ("t0", "<f8"), ("t", "<f8")], align=True)

TR_ABSORBED, TR_ESCAPED, TR_PANEL, TR_ALIVE = 1, 2, 4, 8
COINC_KILL_MULTIPLES, COINC_TAKE_ALL_GOODS = 0, 1


class Event(C.Structure):
    _fields_ = [("parn", C.c_int32), ("pann", C.c_int32), ("modn", C.c_int32), ("cryn", C.c_int32),
                ("siten", C.c_int32), ("eventid", C.c_int32), ("t", C.c_double), ("E", C.c_float),
                ("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Tables(C.Structure):
    _fields_ = [("nmat", C.c_int32), ("eminph", C.c_float), ("emax", C.c_float), ("n1d", C.c_int32),
                ("e0", C.c_float * 4), ("idl", C.c_float * 4), ("mu", C.POINTER(C.c_float) * 4),
                ("ncp", C.c_int32), ("ne", C.c_int32), ("idcp", C.c_float * 2), ("ide", C.c_float * 2),
                ("surf", C.POINTER(C.c_float) * 2)]


class Panel(C.Structure):
    _fields_ = [("panel", C.c_int32)] + [(n, C.c_float) for n in (
        "lenx", "leny", "lenz", "modx", "mody", "modz", "mspx", "mspy", "mspz", "crx", "cry", "crz",
        "cspx", "cspy", "cspz", "ox", "oy", "oz", "dirx", "diry", "dirz")] + [
        ("ux", C.c_float * 3), ("uy", C.c_float * 3), ("uz", C.c_float * 3)]


class PanelMaterial(C.Structure):
    _fields_ = [("material", C.c_int32), ("density", C.c_float)]


class Digitizer(C.Structure):
    _fields_ = [("readout_depth", C.c_int32), ("readout_policy", C.c_int32), ("e_threshold", C.c_float),
                ("e_upholder", C.c_float), ("blur_policy", C.c_int32), ("e_ref", C.c_float), ("r_ref", C.c_float),
                ("slope", C.c_float), ("space_blur", C.c_float), ("dead_level", C.c_int32),
                ("dead_type", C.c_int32), ("dead_time_us", C.c_float), ("win_min", C.c_float),
                ("win_max", C.c_float), ("time_blur_fwhm_us", C.c_double), ("coinc_window_us", C.c_double),
                ("coinc_policy", C.c_int32), ("record_rayleigh", C.c_int32), ("noise_rate", C.c_double),
                ("noise_e_mean", C.c_float), ("noise_e_sigma", C.c_float), ("delay_us", C.c_double)]


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("seed", C.c_uint64), ("eabs", C.c_float), ("noncol_sigma", C.c_float),
                ("positron_range", C.c_int32), ("record_hits", C.c_int32), ("batch", C.c_int64),
                ("flags", C.c_int32), ("record_sphere", C.c_float * 4)]


class Source(C.Structure):
    _fields_ = [("natom", C.c_double), ("isotope", C.c_int32), ("shape", C.c_int32), ("coef", C.c_float * 6)]


class Isotope(C.Structure):
    _fields_ = [("half_life_s", C.c_float), ("beta_ratio", C.c_float), ("coef", C.c_float * 8)]


class Sink(C.Structure):
    _fields_ = [("singles", C.c_void_p), ("singles_cap", C.c_int64), ("n_singles", C.c_int64),
                ("coinc", C.c_void_p), ("coinc_cap", C.c_int64), ("n_coinc", C.c_int64),
                ("adder", C.c_void_p), ("adder_cap", C.c_int64), ("n_adder", C.c_int64),
                ("hit_id", C.c_void_p), ("hit_f", C.c_void_p), ("hits_cap", C.c_int64), ("n_hits", C.c_int64),
                ("trace", C.c_void_p), ("trace_cap", C.c_int64), ("n_trace", C.c_int64),
                ("out_dir", C.c_char_p), ("job_counts", C.c_void_p),
                ("psf_source", C.c_void_p), ("psf_source_cap", C.c_int64), ("n_psf_source", C.c_int64),
                ("psf_phantom", C.c_void_p), ("psf_phantom_cap", C.c_int64), ("n_psf_phantom", C.c_int64),
                ("out_append", C.c_int32),
                ("delayed", C.c_void_p), ("delayed_cap", C.c_int64), ("n_delayed", C.c_int64),
                ("async_copy", C.c_int32)]


PSF_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4"),
                      ("E", "<f4"), ("parn", "<i4"), ("eventid", "<i4"), ("pad", "<i4"), ("t", "<f8")])
assert PSF_DTYPE.itemsize == 48

# list-mode coincidence record (gpet_lm_record, 32 B)
LM_DTYPE = np.dtype([("crystal_a", "<u4"), ("crystal_b", "<u4"), ("t_us", "<f8"), ("dt_ps", "<f4"),
                     ("e_a_kev", "<f4"), ("e_b_kev", "<f4"), ("flags", "<u4")])
assert LM_DTYPE.itemsize == 32
LM_TRUE, LM_DELAYED, LM_BACKGROUND = 1, 2, 4


class Launch(C.Structure):
    _fields_ = [("block", C.c_int32 * 2), ("wpe", C.c_int32 * 2), ("blocks_per_cu", C.c_int32 * 2),
                ("batch", C.c_int64), ("transit", C.c_int32 * 4)]

    def as_dict(self):
        return {"block": list(self.block), "wpe": list(self.wpe), "blocks_per_cu": list(self.blocks_per_cu),
                "batch": int(self.batch), "transit": list(self.transit)}


class SurrogateConfig(C.Structure):
    _fields_ = [("steps", C.c_int64), ("batch", C.c_int32), ("lr", C.c_float), ("seed", C.c_uint64),
                ("reinit", C.c_int32)]


class SurrogateReport(C.Structure):
    _fields_ = [("records", C.c_int64), ("steps", C.c_int64), ("seconds", C.c_double), ("flops", C.c_double),
                ("loss_first", C.c_float), ("loss_last", C.c_float)]


SURR_NX = 6
SURR_NY = 5
SURR_NH = 190


class Job(C.Structure):
    _fields_ = [("phantom", C.c_int32), ("first_source", C.c_int32), ("n_sources", C.c_int32),
                ("reserved", C.c_int32), ("t0", C.c_double), ("t1", C.c_double), ("n_decays", C.c_int64)]


class Tallies(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "n_histories", "n_photons", "n_panel_entries", "n_hits", "n_adder", "n_threshold", "n_deadtime",
        "n_singles", "n_coinc", "n_coinc_true", "dropped_events", "dropped_hits", "majorant_violations",
        "phantom_steps", "range_steps", "panel_steps")] + [
        ("crystal_singles", C.c_void_p), ("n_crystals", C.c_int32),
        ("energy_hist", C.c_void_p), ("n_ebins", C.c_int32), ("e_lo", C.c_float), ("e_hi", C.c_float),
        ("sinogram", C.c_void_p), ("n_s", C.c_int32), ("n_phi", C.c_int32), ("s_max", C.c_float),
        ("seconds_transport", C.c_double), ("seconds_digitizer", C.c_double),
        ("kernel_seconds", C.c_double * 2), ("kernel_launches", C.c_int64 * 2), ("kernel_histories", C.c_int64),
        ("launches", C.c_int64),
        ("launch", Launch), ("graph", C.c_int32), ("n_jobs", C.c_int32), ("seconds_tune", C.c_double),
        ("n_surrogate", C.c_int64), ("seconds_surrogate", C.c_double), ("n_noise", C.c_int64),
        ("n_delayed", C.c_int64), ("sinogram_delayed", C.c_void_p), ("range_jumps", C.c_int64)]

FLAG_NO_GRAPH = 1
FLAG_AUTOTUNE = 2
FLAG_PSF_SOURCE = 4
FLAG_PSF_PHANTOM = 8
FLAG_LISTMODE = 16


class Input(C.Structure):
    _fields_ = [("device", C.c_int32), ("noncol_sigma", C.c_float), ("dims", C.c_int32 * 3),
                ("offset", C.c_float * 3), ("size", C.c_float * 3), ("mat_file", C.c_char * GPET_PATH_MAX),
                ("den_file", C.c_char * GPET_PATH_MAX), ("n_histories", C.c_int64), ("use_psf", C.c_int32),
                ("source_file", C.c_char * GPET_PATH_MAX), ("ptype", C.c_int32), ("prange", C.c_int32),
                ("t0", C.c_float), ("t1", C.c_float), ("record_sphere", C.c_float * 4), ("eabs", C.c_float),
                ("geo_file", C.c_char * GPET_PATH_MAX), ("n_surfaces", C.c_int32),
                ("surfaces", C.c_float * (10 * GPET_MAXSURFACE)), ("readout_depth", C.c_int32),
                ("readout_policy", C.c_int32), ("e_threshold", C.c_float), ("blur_policy", C.c_int32),
                ("e_ref", C.c_float), ("r_ref", C.c_float), ("slope", C.c_float), ("space_blur", C.c_float),
                ("dead_level", C.c_int32), ("dead_type", C.c_int32), ("dead_time", C.c_float),
                ("win_min", C.c_float), ("win_max", C.c_float), ("time_blur_fwhm_us", C.c_double),
                ("coinc_window_us", C.c_double), ("coinc_policy", C.c_int32), ("seed", C.c_uint64),
                ("noise_rate", C.c_double), ("noise_e_mean", C.c_float), ("noise_e_sigma", C.c_float),
                ("delay_us", C.c_double)]


P = C.POINTER
_VP = C.c_void_p

# name -> (restype, argtypes) for the functions shared by both implementations
CTX_FUNCS = {
    "create": (C.c_int32, [P(C.c_void_p), P(Options)]),
    "destroy": (C.c_int32, [_VP]),
    "last_error": (C.c_char_p, [_VP]),
    "set_tables": (C.c_int32, [_VP, P(Tables)]),
    "set_phantom": (C.c_int32, [_VP, _VP, _VP, P(C.c_int32), P(C.c_float), P(C.c_float)]),
    "set_phantom_slot": (C.c_int32, [_VP, C.c_int32, _VP, _VP, P(C.c_int32), P(C.c_float), P(C.c_float)]),
    "set_detector": (C.c_int32, [_VP, P(Panel), C.c_int32, P(PanelMaterial), _VP, C.c_int32]),
    "set_digitizer": (C.c_int32, [_VP, P(Digitizer)]),
    "run_source": (C.c_int32, [_VP, P(Source), C.c_int32, P(Isotope), C.c_int32, C.c_double, C.c_double,
                               C.c_int64, P(Sink)]),
    "run_psf": (C.c_int32, [_VP, _VP, C.c_int64, C.c_int32, P(Sink)]),
    "run_batch": (C.c_int32, [_VP, P(Job), C.c_int32, P(Source), C.c_int32, P(Isotope), C.c_int32, P(Sink)]),
    "set_launch": (C.c_int32, [_VP, P(Launch)]),
    "get_launch": (C.c_int32, [_VP, P(Launch)]),
    "surrogate_harvest": (C.c_int32, [_VP, P(Source), C.c_int32, P(Isotope), C.c_int32, C.c_double, C.c_double,
                                      C.c_int64, C.c_int64]),
    "surrogate_train": (C.c_int32, [_VP, P(SurrogateConfig), P(SurrogateReport)]),
    "surrogate_enable": (C.c_int32, [_VP, C.c_int32]),
    "surrogate_save": (C.c_int32, [_VP, C.c_char_p]),
    "surrogate_load": (C.c_int32, [_VP, C.c_char_p]),
    "surrogate_get_params": (C.c_int32, [_VP, _VP, C.c_int64, P(C.c_int64)]),
    "surrogate_set_params": (C.c_int32, [_VP, _VP, C.c_int64]),
    "surrogate_forward": (C.c_int32, [_VP, _VP, C.c_int64, _VP]),
    "surrogate_records": (C.c_int32, [_VP, _VP, C.c_int64, P(C.c_int64)]),
    "surrogate_grad": (C.c_int32, [_VP, _VP, _VP, C.c_int64, P(C.c_float), _VP]),
    "get_tallies": (C.c_int32, [_VP, P(Tallies)]),
    "get_majorant": (C.c_int32, [_VP, C.c_int32, _VP, P(C.c_float), P(C.c_float)]),
}

HOST_FUNCS = {
    "gpet_tables_load": (C.c_int32, [C.c_char_p, P(Tables)]),
    "gpet_tables_free": (None, [P(Tables)]),
    "gpet_read_geometry": (C.c_int32, [C.c_char_p, P(Panel), C.c_int32, P(C.c_int32), P(PanelMaterial)]),
    "gpet_read_sources": (C.c_int32, [C.c_char_p, P(Source), C.c_int32, P(C.c_int32)]),
    "gpet_read_isotopes": (C.c_int32, [C.c_char_p, P(Isotope), C.c_int32, P(C.c_int32)]),
    "gpet_read_psf": (C.c_int32, [C.c_char_p, _VP, C.c_int64, P(C.c_int64)]),
    "gpet_panel_counts": (C.c_int32, [P(Panel), P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
    "gpet_status_string": (C.c_char_p, [C.c_int32]),
    "gpet_read_input": (C.c_int32, [C.c_char_p, P(Input)]),
    "gpet_input_digitizer": (None, [P(Input), P(Digitizer)]),
    "gpet_write_events": (C.c_int32, [C.c_char_p, _VP, C.c_int64, C.c_int32]),
    "gpet_write_hits": (C.c_int32, [C.c_char_p, C.c_char_p, _VP, _VP, C.c_int64, C.c_int32]),
    "gpet_write_psf": (C.c_int32, [C.c_char_p, C.c_char_p, _VP, C.c_int64, C.c_int32]),
    "gpet_listmode_records": (C.c_int32, [_VP, C.c_int64, C.c_int32, C.c_int32, C.c_uint32, _VP]),
    "gpet_write_listmode": (C.c_int32, [C.c_char_p, _VP, C.c_int64, _VP, C.c_int64, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_int32]),
    "gpet_write_sinogram": (C.c_int32, [C.c_char_p, _VP, C.c_int32, C.c_int32, C.c_float]),
    "gpet_build_majorant": (C.c_int32, [P(Tables), _VP, _VP, P(C.c_float), P(C.c_float)]),
}


def bind(lib: C.CDLL, name: str, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f
