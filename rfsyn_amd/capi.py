"""ctypes mirror of the engine/fdtd C API (include/*.h).

The structs and prototypes here follow include/grid.h, cpml.h, materials.h,
source.h, probe.h, simulate.h and fdtd_native.h, which restate the
reference headers engine/fdtd/{grid,cpml,materials,source,probe,simulate}.h.
Any library exporting that C ABI can be bound with ``bind(cdll)``; the
product library is loaded by ``rfsyn_amd.load_library()``.
"""
import ctypes as C

FIELD_EX, FIELD_EY, FIELD_EZ, FIELD_HX, FIELD_HY, FIELD_HZ = range(6)
SRC_SOFT, SRC_HARD = 0, 1
FDTD_FP64, FDTD_FP32 = 0, 1
MATERIALS_MAX_POLES = 4

_dp = C.POINTER(C.c_double)


class Grid(C.Structure):
    """fdtd_grid_t (include/grid.h; reference engine/fdtd/grid.h:6-13).
    The trailing ``dev`` pointer is library-private."""
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t),
                ("dx", C.c_double), ("dt", C.c_double),
                ("periodic_axes", C.c_uint),
                ("ex", _dp), ("ey", _dp), ("ez", _dp),
                ("hx", _dp), ("hy", _dp), ("hz", _dp),
                ("dev", C.c_void_p)]


_PROFILE_NAMES = [f"{k}_{mh}_{ax}" for mh in ("m", "h") for ax in "xyz"
                  for k in ("kappa_inv", "b", "a")]
_PSI_NAMES = ["psi_hx_y", "psi_hx_z", "psi_hy_z", "psi_hy_x", "psi_hz_x",
              "psi_hz_y", "psi_ex_y", "psi_ex_z", "psi_ey_z", "psi_ey_x",
              "psi_ez_x", "psi_ez_y"]


class Cpml(C.Structure):
    """struct cpml (include/cpml.h; reference engine/fdtd/cpml.h:8-26).
    Fields after the psi pointers are library-private."""
    _fields_ = ([("npml", C.c_size_t), ("dx", C.c_double)]
                + [(n, _dp) for n in _PROFILE_NAMES]
                + [(n, _dp) for n in _PSI_NAMES]
                + [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t),
                   ("periodic_axes", C.c_uint), ("dt", C.c_double),
                   ("dev", C.c_void_p)])


class CpmlParams(C.Structure):
    _fields_ = [("npml", C.c_size_t), ("m", C.c_double), ("ma", C.c_double),
                ("kappa_max", C.c_double), ("alpha_max", C.c_double),
                ("r0", C.c_double)]


class MaterialCoeff(C.Structure):
    _fields_ = [("Ca", C.c_double), ("Cb", C.c_double)]


class MaterialTensor(C.Structure):
    _fields_ = [("m11", C.c_double), ("m12", C.c_double), ("m22", C.c_double),
                ("scale", C.c_double)]


class MaterialDispersive(C.Structure):
    _fields_ = [("eps_inf", C.c_double), ("sigma", C.c_double),
                ("n_drude", C.c_size_t),
                ("drude_wp", C.c_double * MATERIALS_MAX_POLES),
                ("drude_gamma", C.c_double * MATERIALS_MAX_POLES),
                ("n_lorentz", C.c_size_t),
                ("lorentz_deps", C.c_double * MATERIALS_MAX_POLES),
                ("lorentz_w0", C.c_double * MATERIALS_MAX_POLES),
                ("lorentz_delta", C.c_double * MATERIALS_MAX_POLES)]


WAVEFORM = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class Source(C.Structure):
    _fields_ = [("i", C.c_size_t), ("j", C.c_size_t), ("k", C.c_size_t),
                ("component", C.c_int), ("mode", C.c_int),
                ("amplitude", C.c_double), ("waveform", C.c_void_p),
                ("params", C.c_void_p)]


class GaussDerivParams(C.Structure):
    _fields_ = [("t0", C.c_double), ("tau", C.c_double)]


class RickerParams(C.Structure):
    _fields_ = [("f0_hz", C.c_double), ("t0", C.c_double)]


class ModGaussParams(C.Structure):
    _fields_ = [("freq_hz", C.c_double), ("t0", C.c_double), ("tau", C.c_double)]


class ProbeTD(C.Structure):
    _fields_ = [("i", C.c_size_t), ("j", C.c_size_t), ("k", C.c_size_t),
                ("component", C.c_int), ("buffer", _dp),
                ("n_steps", C.c_size_t), ("next_index", C.c_size_t)]


class ProbeDFT(C.Structure):
    _fields_ = [("i", C.c_size_t), ("j", C.c_size_t), ("k", C.c_size_t),
                ("component", C.c_int), ("freqs_hz", _dp),
                ("n_freqs", C.c_size_t), ("acc_re", _dp), ("acc_im", _dp)]


class SimConfig(C.Structure):
    """fdtd_sim_config_t (reference engine/fdtd/simulate.h:9-23)."""
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t),
                ("dx", C.c_double), ("courant_safety", C.c_double),
                ("n_steps", C.c_size_t), ("npml", C.c_size_t),
                ("src_i", C.c_size_t), ("src_j", C.c_size_t), ("src_k", C.c_size_t),
                ("src_component", C.c_int),
                ("t0_over_dt", C.c_double), ("tau_over_dt", C.c_double),
                ("probe_i", C.c_size_t), ("probe_j", C.c_size_t), ("probe_k", C.c_size_t),
                ("probe_component", C.c_int)]


class Shape(C.Structure):
    """fdtd_shape_t (include/fdtd_scene.h)."""
    _fields_ = [("kind", C.c_int), ("eps_r", C.c_double), ("p", C.c_double * 3), ("radius", C.c_double),
                ("normal", C.c_double * 3), ("half_thickness", C.c_double), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3)]


class SceneOptions(C.Structure):
    _fields_ = [("supersample", C.c_int), ("keepout", C.c_size_t)]


class SceneStats(C.Structure):
    _fields_ = [("nodes_painted", C.c_size_t), ("override_edges", C.c_size_t),
                ("tensor_edges", C.c_size_t), ("scalar_mixed_edges", C.c_size_t)]


SHAPE_SPHERE, SHAPE_SLAB, SHAPE_BOX = 0, 1, 2


class SimOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("dft_compensated", C.c_int),
                ("dft_block", C.c_size_t)]


class SurConfig(C.Structure):
    """fdtd_sur_config_t (include/fdtd_surrogate.h)."""
    _fields_ = [("grid", C.c_int), ("in_channels", C.c_int), ("channels", C.c_int * 4), ("coords", C.c_int),
                ("width", C.c_int), ("outputs", C.c_int), ("queries", C.c_int), ("max_batch", C.c_int)]


_fp = C.POINTER(C.c_float)
_sz = C.c_size_t
_vp = C.c_void_p
_gp = C.POINTER(Grid)
_cp = C.POINTER(Cpml)

# name -> (restype, argtypes); the reference API first, then extensions.
REFERENCE_API = {
    "fdtd_grid_create": (_gp, [_sz, _sz, _sz, C.c_double, C.c_double]),
    "fdtd_grid_set_periodic": (C.c_int, [_gp, C.c_uint]),
    "fdtd_grid_destroy": (None, [_gp]),
    "fdtd_update_h": (None, [_gp]),
    "fdtd_update_e": (None, [_gp]),
    "cpml_default_params": (CpmlParams, [_sz]),
    "cpml_create": (_cp, [_gp, CpmlParams]),
    "cpml_destroy": (None, [_cp]),
    "fdtd_update_h_cpml": (None, [_gp, _cp]),
    "fdtd_update_e_cpml": (None, [_gp, _cp]),
    "materials_create": (_vp, [_gp]),
    "materials_destroy": (None, [_vp]),
    "materials_register": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(C.c_uint8)]),
    "materials_assign_box": (C.c_int, [_vp, _gp, _sz, _sz, _sz, _sz, _sz, _sz, C.c_uint8]),
    "materials_set_box": (C.c_int, [_vp, _gp, _sz, _sz, _sz, _sz, _sz, _sz, C.c_double, C.c_double]),
    "materials_set_component": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz, C.c_double, C.c_double]),
    "materials_clear_component": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz]),
    "materials_get_coefficients": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz, C.POINTER(MaterialCoeff)]),
    "fdtd_update_e_materials": (None, [_gp, _vp]),
    "fdtd_update_e_cpml_materials": (None, [_gp, _cp, _vp]),
    "materials_set_component_tensor": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz, C.c_double, C.c_double, C.c_double]),
    "materials_clear_component_tensor": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz]),
    "materials_get_tensor": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz, C.POINTER(MaterialTensor)]),
    "source_waveform_gaussian_derivative": (C.c_double, [C.c_double, _vp]),
    "source_waveform_ricker": (C.c_double, [C.c_double, _vp]),
    "source_waveform_modulated_gaussian": (C.c_double, [C.c_double, _vp]),
    "source_make_point": (Source, [_sz, _sz, _sz, C.c_int, C.c_int, C.c_double, _vp, _vp]),
    "source_apply": (None, [C.POINTER(Source), _gp, C.c_double]),
    "probe_td_create": (C.c_int, [C.POINTER(ProbeTD), _sz, _sz, _sz, C.c_int, _sz]),
    "probe_td_destroy": (None, [C.POINTER(ProbeTD)]),
    "probe_td_record": (None, [C.POINTER(ProbeTD), _gp]),
    "probe_dft_create": (C.c_int, [C.POINTER(ProbeDFT), _sz, _sz, _sz, C.c_int, _dp, _sz]),
    "probe_dft_destroy": (None, [C.POINTER(ProbeDFT)]),
    "probe_dft_record": (None, [C.POINTER(ProbeDFT), _gp, C.c_double, C.c_double]),
    "probe_dft_result": (None, [C.POINTER(ProbeDFT), _sz, _dp, _dp]),
    "fdtd_simulate_echo": (C.c_int, [C.POINTER(SimConfig), _dp, _dp]),
    "fdtd_simulate_echo_with_setup": (C.c_int, [C.POINTER(SimConfig), _dp, _dp, _vp, _vp]),
}

EXTENSION_API = {
    "materials_set_component_tensor3": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz, _dp]),
    "materials_get_tensor3": (C.c_int, [_vp, _gp, C.c_int, _sz, _sz, _sz, _dp]),
    "materials_register_dispersive": (C.c_int, [_vp, C.POINTER(MaterialDispersive), C.POINTER(C.c_uint8)]),
    "materials_table_count": (_sz, [_vp]),
    "materials_get_dispersive": (C.c_int, [_vp, C.c_uint8, C.POINTER(MaterialDispersive)]),
    "materials_override_count": (_sz, [_vp]),
    "materials_tensor_count": (_sz, [_vp]),
    "fdtd_sim_create": (_vp, [_gp, _cp, _vp, C.POINTER(SimOptions)]),
    "fdtd_sim_destroy": (None, [_vp]),
    "fdtd_sim_add_point_source": (C.c_int, [_vp, C.POINTER(Source)]),
    "fdtd_sim_add_plane_source": (C.c_int, [_vp, C.c_int, C.c_int, _sz, C.c_int, C.c_int, C.c_double, _vp, _vp]),
    "fdtd_sim_add_probe": (C.c_int, [_vp, C.c_int, _sz, _sz, _sz, _sz, _dp, _sz]),
    "fdtd_sim_add_plane_monitor": (C.c_int, [_vp, C.c_int, _sz, _dp, _sz]),
    "fdtd_sim_run": (C.c_int, [_vp, _sz]),
    "fdtd_sim_steps_done": (_sz, [_vp]),
    "fdtd_sim_streams": (C.c_int, [_vp]),
    "fdtd_sim_probe_td": (C.c_long, [_vp, C.c_int, _dp, _sz]),
    "fdtd_sim_probe_iq": (C.c_int, [_vp, C.c_int, _sz, C.c_int, C.c_float, C.c_uint64, C.c_uint64,
                                    C.POINTER(C.c_float)]),
    "fdtd_sim_probe_iq_noalloc": (C.c_int, [_vp, C.c_int, _sz, C.c_int, C.c_float, C.c_uint64, C.c_uint64,
                                            C.POINTER(C.c_float)]),
    "fdtd_write_iq": (C.c_int, [C.c_char_p, C.POINTER(C.c_float), _sz]),
    "fdtd_write_transfer": (C.c_int, [C.c_char_p, _dp, _dp, _sz, _sz]),
    "fdtd_read_transfer": (C.c_int, [C.c_char_p, _dp, _dp, _sz, _sz]),
    "fdtd_simulate_echo_iq": (C.c_int, [_vp, _vp, C.POINTER(C.c_float), _dp]),
    "oracle_awgn_iq": (None, [C.POINTER(C.c_float), _sz, C.c_float, C.c_uint64, C.c_uint64]),
    "fdtd_sim_probe_dft": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "fdtd_sim_plane_dft": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "fdtd_sim_download": (C.c_int, [_vp, _gp]),
    "fdtd_sim_upload": (C.c_int, [_vp, _gp]),
    "fdtd_sim_energy": (C.c_int, [_vp, _dp, _dp]),
    "fdtd_sim_device_bytes": (_sz, [_vp]),
    "fdtd_sim_set_profiling": (C.c_int, [_vp, C.c_int]),
    "fdtd_sim_kernel_times": (C.c_int, [_vp, _dp, _dp, C.POINTER(C.c_long)]),
    "fdtd_sim_kernel_times3": (C.c_int, [_vp, _dp, C.POINTER(C.c_long)]),
    "fdtd_sim_energy_audit": (C.c_int, [_vp, _dp, C.POINTER(_sz)]),
    "fdtd_sim_run_many": (C.c_int, [C.POINTER(_vp), _sz, _sz]),
    "fdtd_sim_dft_block": (C.c_int, [_vp]),
    "fdtd_sim_leap_rows": (C.c_int, [_vp]),
    "fdtd_sim_fields_finite": (C.c_int, [_vp]),
    "fdtd_gpu_errors": (C.c_int, []),
    "fdtd_batch_fields_finite": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "fdtd_batch_create": (_vp, [C.POINTER(_vp), _sz]),
    "fdtd_batch_run": (C.c_int, [_vp, _sz]),
    "fdtd_batch_destroy": (None, [_vp]),
    "fdtd_batch_size": (_sz, [_vp]),
    "fdtd_batch_launches": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fdtd_sim_autotune": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "fdtd_sim_select_path": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "fdtd_sur_create": (_vp, [C.POINTER(SurConfig), C.c_uint64]),
    "fdtd_sur_destroy": (None, [_vp]),
    "fdtd_sur_n_params": (_sz, [_vp]),
    "fdtd_sur_get_params": (C.c_int, [_vp, _fp]),
    "fdtd_sur_set_params": (C.c_int, [_vp, _fp]),
    "fdtd_sur_predict": (C.c_int, [_vp, _fp, _fp, C.c_int, _fp]),
    "fdtd_sur_gradient": (C.c_int, [_vp, _fp, _fp, _fp, C.c_int, _fp, _dp]),
    "fdtd_sur_train_step": (C.c_int, [_vp, _fp, _fp, _fp, C.c_int, C.c_float, _dp]),
    "fdtd_sim_add_flux_plane": (C.c_int, [_vp, C.c_int, _sz]),
    "fdtd_sim_flux": (C.c_int, [_vp, C.c_int, _dp]),
    "fdtd_sim_fused": (C.c_int, [_vp]),
    "fdtd_sim_fused_box": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fdtd_last_error": (C.c_char_p, []),
    "fdtd_scene_paint": (C.c_int, [_vp, _gp, C.POINTER(Shape), _sz, C.POINTER(SceneOptions),
                                   C.POINTER(SceneStats)]),
}


def bind(lib, names=None):
    """Declare prototypes on a ctypes.CDLL for every symbol it exports."""
    table = dict(REFERENCE_API)
    table.update(EXTENSION_API)
    for name, (res, args) in table.items():
        if names is not None and name not in names:
            continue
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


def waveform_ptr(lib, name):
    """Address of a waveform function exported by ``lib``."""
    return C.cast(getattr(lib, name), C.c_void_p).value
