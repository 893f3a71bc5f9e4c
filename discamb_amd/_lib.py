"""ctypes binding of include/discamb_amd.h (libdiscamb_amd.so, built in-tree by build())."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiscamb_amd.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_dbl_p = ctypes.POINTER(ctypes.c_double)
c_u8_p = ctypes.POINTER(ctypes.c_ubyte)


class DhModelDesc(ctypes.Structure):
    _fields_ = [
        ("cell", ctypes.c_double * 6),
        ("n_symops", ctypes.c_int),
        ("symop_rot", c_int_p),
        ("symop_trans", c_dbl_p),
        ("xyz_convention", ctypes.c_int),
        ("adp_convention", ctypes.c_int),
        ("d_xyz_convention", ctypes.c_int),
        ("d_adp_convention", ctypes.c_int),
        ("n_wfn", ctypes.c_int),
        ("wfn_orbital_offset", c_int_p),
        ("orbital_term_offset", c_int_p),
        ("orbital_nl", c_int_p),
        ("term_coeff", c_dbl_p),
        ("term_exp", c_dbl_p),
        ("term_pow", c_int_p),
        ("orbital_occupancy", c_int_p),
        ("orbital_kind", c_int_p),
        ("def_val_exponent", c_dbl_p),
        ("def_val_pow_offset", c_int_p),
        ("def_val_pow", c_int_p),
        ("anomalous", c_dbl_p),
        ("n_types", ctypes.c_int),
        ("plm_n_l", c_int_p),
        ("plm", c_dbl_p),
        ("p_val", c_dbl_p),
        ("kappa", c_dbl_p),
        ("kappa_prime", c_dbl_p),
        ("n_atoms", ctypes.c_int),
        ("atom_to_wfn", c_int_p),
        ("atom_to_type", c_int_p),
        ("atom_multiplicity", c_dbl_p),
        ("wfn_label", ctypes.POINTER(ctypes.c_char_p)),
        ("wfn_atomic_number", c_int_p),
    ]


class DhAtoms(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int),
        ("xyz", c_dbl_p),
        ("adp_size", c_int_p),
        ("adp", c_dbl_p),
        ("occupancy", c_dbl_p),
        ("lcs", c_dbl_p),
        ("anomalous", c_dbl_p),
        ("include", c_u8_p),
    ]


class DhAtomsDev(ctypes.Structure):
    """dh_atoms_dev: device addresses (integers) of the per-atom arrays."""
    _fields_ = [
        ("n_atoms", ctypes.c_int),
        ("xyz", ctypes.c_void_p),
        ("adp_size", ctypes.c_void_p),
        ("adp", ctypes.c_void_p),
        ("occupancy", ctypes.c_void_p),
        ("lcs", ctypes.c_void_p),
        ("anomalous", ctypes.c_void_p),
        ("include", ctypes.c_void_p),
    ]


class DhAtomRef(ctypes.Structure):
    _fields_ = [("atom", ctypes.c_int), ("rot", ctypes.c_int * 9), ("trans", ctypes.c_double * 3)]


class DhLcsDef(ctypes.Structure):
    _fields_ = [("central", DhAtomRef), ("direction1_type", ctypes.c_int), ("direction2_type", ctypes.c_int),
                ("n_ref1", ctypes.c_int), ("n_ref2", ctypes.c_int), ("n_chirality", ctypes.c_int),
                ("ref1", ctypes.POINTER(DhAtomRef)), ("ref2", ctypes.POINTER(DhAtomRef)),
                ("chirality", ctypes.POINTER(DhAtomRef)), ("coordinate_1", ctypes.c_int),
                ("coordinate_2", ctypes.c_int), ("is_r", ctypes.c_int)]


# int (*)(void* user, int hkl_index, const double* f, const double* row)
PER_HKL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_dbl_p)
JAC_PLANES = 10

# every symbol include/discamb_amd.h declares, with (restype, argtypes)
SIGNATURES = {
    "dh_last_error": (ctypes.c_char_p, []),
    "dh_version": (ctypes.c_char_p, []),
    "dh_device_count": (ctypes.c_int, [c_int_p]),
    "dh_hc_create": (ctypes.c_int, [ctypes.POINTER(DhModelDesc), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "dh_hc_destroy": (None, [ctypes.c_void_p]),
    "dh_hc_set_calculation_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "dh_hc_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dh_hc_set_electron_scattering": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dh_hc_set_iam_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "dh_hc_set_options": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint]),
    "dh_hc_sf_and_derivs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhAtoms), ctypes.c_int, c_int_p, c_dbl_p,
                                           ctypes.c_uint, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]),
    "dh_hc_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhAtoms), ctypes.c_int, c_int_p, c_dbl_p]),
    "dh_hc_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint]),
    "dh_hc_fetch": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]),
    "dh_hc_set_type_parameters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]),
    "dh_hc_fetch_type_derivatives": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]),
    "dh_hc_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, c_int_p]),
    "dh_hc_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dh_hc_last_timing": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, c_int_p, c_int_p, c_int_p]),
    "dh_hc_engine_info": (ctypes.c_int, [ctypes.c_void_p, c_int_p, c_int_p, c_int_p, ctypes.c_char_p]),
    "dh_hc_set_reflections": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p]),
    "dh_hc_sf_and_derivs_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhAtomsDev), ctypes.c_void_p,
                                               ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]),
    "dh_hc_sf_and_derivs_mp_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhAtomsDev), ctypes.c_void_p,
                                                  ctypes.c_uint] + [ctypes.c_void_p] * 9),
    "dh_hc_sf_and_derivs_dev_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhAtomsDev), ctypes.c_void_p,
                                                       ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dh_hc_per_hkl": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhAtoms), ctypes.c_int, c_int_p, PER_HKL_FN,
                                     ctypes.c_void_p]),
    "dh_hc_jacobian_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DhAtoms), ctypes.c_int, c_int_p]),
    "dh_hc_jacobian_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_dbl_p, c_dbl_p]),
    "dh_hc_jacobian_rows_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]),
    "dh_hc_form_factors": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, c_u8_p, ctypes.c_int, c_int_p, c_dbl_p]),
    "dh_hc_form_factors_cart": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, c_u8_p, ctypes.c_int, c_dbl_p, c_dbl_p]),
    "dh_lcs_compute": (ctypes.c_int, [c_dbl_p, ctypes.c_int, c_dbl_p, ctypes.c_int, ctypes.POINTER(DhLcsDef), c_dbl_p]),
    "dh_hc_set_lcs_definitions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(DhLcsDef), ctypes.c_int]),
    "dh_hc_last_lcs": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p]),
    "dh_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "dh_batch_destroy": (None, [ctypes.c_void_p]),
    "dh_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_int]),
    "dh_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, c_int_p]),
    "dh_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dh_batch_last_timing": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, c_int_p, c_int_p, c_int_p]),
}

_lib = None


def load():
    """Load the HIP library.  There is no CPU fallback: a missing library is an error."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("discamb_amd: %s is not built; run __graft_entry__.build() (make -C discamb_amd)"
                           % LIB_PATH)
    # PyTorch-ROCm wheels carry their own libamdhip64 / libhsa-runtime64, which torch's libraries
    # name without the version suffix.  Loaded after this library, torch would map a second HIP and
    # HSA runtime into the process, and that second runtime finds no GPU.  Importing torch first
    # lets this library's libamdhip64.so.7 bind to torch's copy, so the process has one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class DhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("discamb_amd error %d: %s" % (code, msg))
        self.code = code


def check(code):
    if code != 0:
        raise DhError(code, load().dh_last_error().decode())


def ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))
