"""ctypes mirrors of the C-ABI in include/pureibm_gpu.h and loaders for the
in-tree shared libraries.

``load_gpu()`` returns the HIP library (libpureibm_gpu.so) and raises if it
is missing: there is no CPU fallback for the GPU path.  ``load_host()``
returns the Fortran host library (libpib_host.so: namelists, configuration,
point set).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

PIB_OK = 0
PIB_NERR_DEFAULT = 100  # drag-error history length at creation (pureibm_gpu.h)
PIB_NERR_MAX = 4096
PIB_ERR_ARG = -1
PIB_ERR_UNSUPPORTED = -2
PIB_ERR_HIP = -3
PIB_ERR_STATE = -4
PIB_ERR_BLOWUP = -5
PIB_ERR_PIVOT = -6
PIB_ERR_MOTION = -7
PIB_ERR_CAPACITY = -8
PIB_TRIDIAG_NOALLOC_MAX = 1 << 20  # rows of the no-alloc tridiagonal workspace

PIB_TUNE_VELSTEP = 0
PIB_TUNE_GRAPH = 1
PIB_TUNE_STREAMS = 4
PIB_TUNE_NL = 5
PIB_TUNE_FFT = 2

PIB_FP64 = 0
PIB_FP32_BULK = 1

PIB_SCHEME_CENTER = 0
PIB_SCHEME_UPWIND = 1
PIB_LIMITER_NONE = 0
PIB_LIMITER_MINMOD = 1

FIELDS = {
    "u": 0, "p": 1, "nl": 2, "onl": 3, "ff": 4, "ubcp": 5, "pbcp": 6,
    "nlbcp": 7, "onlbcp": 8, "diffn": 9, "ppr": 10, "fr": 11, "omega": 12,
    "mask": 13, "phi": 14, "nlphi": 15, "onlphi": 16, "ffphi": 17,
    "au": 18, "ap": 19, "anl": 20,
}
COMPLEX_FIELDS = {"u", "p", "nl", "onl", "ff", "phi", "nlphi", "onlphi", "ffphi", "au", "ap", "anl"}


class PibConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("my", ctypes.c_int), ("mz", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("dx", ctypes.c_double), ("vis", ctypes.c_double), ("cfl_max", ctypes.c_double),
        ("rk3", ctypes.c_int), ("adaptive", ctypes.c_int), ("discr_scheme", ctypes.c_int),
        ("limiter", ctypes.c_int), ("revernorm", ctypes.c_int), ("dorpr", ctypes.c_int),
        ("dobnd", ctypes.c_int), ("include_frmeanfluid", ctypes.c_int),
        ("set_mpg", ctypes.c_int), ("soft_start", ctypes.c_int),
        ("move_particles", ctypes.c_int), ("reserved_i", ctypes.c_int * 1),
        ("ucharmod", ctypes.c_double), ("char_length", ctypes.c_double),
        ("dia", ctypes.c_double), ("dr", ctypes.c_double), ("voldom", ctypes.c_double),
        ("rhos", ctypes.c_double), ("rhof", ctypes.c_double),
        ("reserved_d", ctypes.c_double * 6),
    ]


_V3 = ctypes.c_double * 3


class PibScalars(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double), ("dt", ctypes.c_double), ("cf", ctypes.c_double),
        ("cforig", ctypes.c_double),
        ("umean", _V3), ("mpg", _V3), ("frmean", _V3), ("frmeanfluid", _V3),
        ("frame_accln", _V3),
        ("ufmean", _V3), ("ufmean_old", _V3), ("ufmean_des", _V3),
        ("usmean", _V3), ("usmean_des", _V3), ("usmean_act", _V3),
        ("dufmeandt", _V3), ("umean_temp", _V3), ("usmean_temp", _V3),
        ("meanslipmod", ctypes.c_double), ("mixmeanslipmod", ctypes.c_double),
        ("maxvolfrac", ctypes.c_double), ("mean_volfrac", ctypes.c_double),
        ("pres_total", _V3), ("visc_total", _V3), ("visc_total_old", _V3),
        ("pres_drag", ctypes.c_double), ("visc_drag", ctypes.c_double),
        ("total_drag", ctypes.c_double), ("norm_drag", ctypes.c_double),
        ("norm_drag_chem", ctypes.c_double),
        ("ferror", ctypes.c_double), ("fold", ctypes.c_double),
        ("total_drag_mpg", ctypes.c_double),
        ("dtnew", ctypes.c_double), ("umax_tmp", ctypes.c_double), ("divmax", ctypes.c_double),
        ("count_solid", ctypes.c_double), ("count_fluid", ctypes.c_double),
        ("ramp_frac_time", ctypes.c_double),
        ("frame_vel", ctypes.c_double * 3), ("frame_pos", ctypes.c_double * 3),
        ("reserved", ctypes.c_double * 9),
    ]

    def as_dict(self):
        out = {}
        for name, typ in self._fields_:
            v = getattr(self, name)
            if name.startswith("reserved"):
                continue
            out[name] = list(v) if hasattr(v, "__len__") else float(v)
        return out


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_PLL = ctypes.POINTER(ctypes.c_longlong)

# entry points shared by the GPU library (pib_) and the oracle (orc_)
_STEP_SIGS = {
    "create": ([ctypes.POINTER(PibConfig), ctypes.POINTER(_VP)], _I),
    "destroy": ([_VP], _I),
    "last_error": ([_VP], ctypes.c_char_p),
    "set_points": ([_VP, _I, _PD, _PD], _I),
    "set_particles": ([_VP, _I, _PD, _PD, _PD, _PD, _PLL, _PI], _I),
    "set_field": ([_VP, _I, _PD], _I),
    "get_field": ([_VP, _I, _PD], _I),
    "set_scalars": ([_VP, ctypes.POINTER(PibScalars)], _I),
    "get_scalars": ([_VP, ctypes.POINTER(PibScalars)], _I),
    "get_ferror_history": ([_VP, _I, _PD, _PD], _I),
    "set_ferror_history": ([_VP, _I, _PD], _I),
    "get_forces": ([_VP, _PD, _PD, _PD, _PD], _I),
    "compute_new_timestep": ([_VP, _I], _I),
    "nonlinear": ([_VP, _I], _I),
    "nonlinear_begin": ([_VP, _I], _I),
    "nonlinear_finish": ([_VP, _I], _I),
    "velstep": ([_VP, _I], _I),
    "step": ([_VP, _I], _I),
    "scalar_set": ([_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double], _I),
    "scalar_get": ([_VP, _PD], _I),
    "scalstep_modes": ([_VP, _I], _I),
    "scalstep": ([_VP, _I], _I),
    "scalar_surface": ([_VP, _PD], _I),
    "scalar_flux": ([_VP, _PD, _PD], _I),
    "scalar_ratios": ([_VP, _PD], _I),
    "adjoint_begin": ([_VP], _I),
    "adjoint_step": ([_VP, _PD], _I),
    "adjoint_record": ([_VP, _I], _I),
    "adjoint_get": ([_VP, _PD, _PD, _PD], _I),
    "adjoint_set_umean": ([_VP, _PD], _I),
    "checkpoint_push": ([_VP], _I),
    "checkpoint_pop": ([_VP], _I),
    "checkpoint_steps": ([_VP, _I], _I),
}

_GPU_ONLY_SIGS = {
    "step_async": ([_VP, _I], _I),
    "sync": ([_VP], _I),
    "device_info": ([ctypes.c_char_p, _I, _PI], _I),
    "ff2rc_planes": ([_VP, _PD, ctypes.c_longlong, _I, _PD, ctypes.c_longlong], _I),
    "ff2cr_planes": ([_VP, _PD, ctypes.c_longlong, _I, _PD, ctypes.c_longlong], _I),
    "tridiag_batched": ([_VP, _I, _I, _PD, _PD, _PD, _PD, _PD, _I], _I),
    "tridiag_batched_noalloc": ([_VP, _I, _I, _VP, _VP, _VP, _VP, _VP, _I, _VP], _I),
    "phase_times": ([_VP, _PD], _I),
    "set_timing": ([_VP, _I], _I),
    "kernel_times": ([_VP, _I, ctypes.c_char_p, _PD, _PI, _PI], _I),
    "set_tuning": ([_VP, _I, _I], _I),
    "get_tuning": ([_VP, _I, _PI], _I),
    "autotune": ([_VP, _I, _I, _PI, _PD], _I),
    "batch_create": ([ctypes.POINTER(_VP), _I, ctypes.POINTER(_VP)], _I),
    "batch_step": ([_VP, _I], _I),
    "batch_sync": ([_VP], _I),
    "batch_destroy": ([_VP], _I),
    "batch_last_error": ([_VP], ctypes.c_char_p),
}



class PibDemConfig(ctypes.Structure):
    """pib_dem_config (include/pureibm_dem.h)."""
    _fields_ = [
        ("nbody", ctypes.c_int), ("only_dem", ctypes.c_int), ("reserved_i", ctypes.c_int * 2),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("xlength", ctypes.c_double), ("ylength", ctypes.c_double), ("zlength", ctypes.c_double),
        ("foffset", ctypes.c_double), ("rhos", ctypes.c_double), ("rhof", ctypes.c_double),
        ("dia_phys", ctypes.c_double), ("coeff_rest", ctypes.c_double), ("kn", ctypes.c_double),
        ("reserved_d", ctypes.c_double * 6),
    ]


class PibDemInfo(ctypes.Structure):
    """pib_dem_info (include/pureibm_dem.h)."""
    _fields_ = [
        ("dtsolid", ctypes.c_double), ("tcoll", ctypes.c_double), ("dtsolid_used", ctypes.c_double),
        ("s_time", ctypes.c_double), ("overlap_max", ctypes.c_double), ("box", ctypes.c_double * 3),
        ("factor", ctypes.c_int), ("ncontacts", ctypes.c_int), ("cells", ctypes.c_int * 3),
        ("substeps", ctypes.c_int), ("rebuilds", ctypes.c_int), ("reserved", ctypes.c_int * 7),
    ]


# DEM entry points shared by the GPU library (pib_dem_) and the oracle (orc_dem_)
_DEM_SIGS = {
    "dem_create": ([ctypes.POINTER(PibDemConfig), _PD, _PD, _PD, ctypes.POINTER(_VP)], _I),
    "dem_destroy": ([_VP], _I),
    "dem_last_error": ([_VP], ctypes.c_char_p),
    "dem_march": ([_VP, _D, _PD, _PD, _PD, _PD, _PD, _PD, _PD], _I),
    "dem_get_state": ([_VP, _PD, _PD, _PD, _PD, ctypes.POINTER(PibDemInfo)], _I),
    "dem_set_state": ([_VP, _PD, _PD], _I),
}

GPU_SYMBOLS = ["pib_" + k for k in list(_STEP_SIGS) + list(_GPU_ONLY_SIGS) + list(_DEM_SIGS)]


def bind(lib, prefix, sigs):
    for name, (args, res) in sigs.items():
        fn = getattr(lib, prefix + name)
        fn.argtypes = args
        fn.restype = res
    return lib


_gpu = None
_host = None


def gpu_lib_path():
    return os.path.join(BUILD, "libpureibm_gpu.so")


def host_lib_path():
    return os.path.join(BUILD, "libpib_host.so")


def load_gpu():
    """The HIP library; raises when it has not been built (no fallback)."""
    global _gpu
    if _gpu is None:
        path = gpu_lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"HIP library {path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        bind(lib, "pib_", _STEP_SIGS)
        bind(lib, "pib_", _GPU_ONLY_SIGS)
        bind(lib, "pib_", _DEM_SIGS)
        _gpu = lib
    return _gpu


_HOST_SIGS = {
    "read": ([ctypes.c_char_p, _I], _I),
    "configure": ([ctypes.c_char_p, _I], _I),
    "finish": ([ctypes.c_longlong, ctypes.POINTER(PibScalars)], _I),
    "counts": ([_PI, _PI, _PI, _PI, _PI], _I),
    "get_config": ([ctypes.POINTER(PibConfig)], _I),
    "get_particles": ([_PD, _PD], _I),
    "get_points": ([_PD, _PD], _I),
    "last_error": ([ctypes.c_char_p, _I], _I),
    "quadgener": ([_D, _D, _I, _PD, _PD], _I),
    "set_option": ([_I, _D], _I),
    "struct_sizes": ([_PI], _I),
    "get_dem_config": ([ctypes.POINTER(PibDemConfig)], _I),
    "get_nsteps_hint": ([], _I),
    "save_restart": ([ctypes.c_char_p, _I, ctypes.POINTER(PibScalars), _I, _PD, _PD, _PD, _I, _I, _I,
                      _PD, _PD, _PD, _PD, _PI], _I),
    "read_restart": ([ctypes.c_char_p, _I, ctypes.POINTER(PibScalars), _I, _PD, _PD, _PD, _I, _I, _I,
                      _PD, _PD, _PD, _PD, _PI], _I),
    "get_scalar": ([_PD], _I),
    "save_scal_restart": ([ctypes.c_char_p, _I, _I, _I, _I, _I, _I, _PD, _PD, _PD, _PD, _PD, _PD], _I),
    "read_scal_restart": ([ctypes.c_char_p, _I, _I, _I, _I, _I, _I, _PD, _PD, _PD, _PD, _PD, _PD], _I),
}


def load_host():
    global _host
    if _host is None:
        path = host_lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"host library {path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        bind(lib, "pib_host_", _HOST_SIGS)
        _host = lib
    return _host


def host_error():
    buf = ctypes.create_string_buffer(600)
    load_host().pib_host_last_error(buf, 600)
    return buf.value.decode(errors="replace")
