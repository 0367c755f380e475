"""ctypes declarations shared by the golden-vector generator and the oracle
unit tests: the reference's pointwise functions (exported by
oracle/_ref/libgamera_ref.so) and the oracle's restatements of them
(oracle/liboracle.so), which take the same C types.
"""
import ctypes

D = ctypes.c_double
PD = ctypes.POINTER(D)
I = ctypes.c_int
B = ctypes.c_bool


class Vec3(ctypes.Structure):
    _fields_ = [("value", D * 3)]


class Primitive(ctypes.Structure):
    _fields_ = [("density", D), ("velocity", Vec3), ("pressure", D)]


class FaceGeometry(ctypes.Structure):
    _fields_ = [("area", D), ("area_vector", Vec3), ("centroid", Vec3),
                ("normal", Vec3), ("tangent1", Vec3), ("tangent2", Vec3)]


class CellGeometry(ctypes.Structure):
    _fields_ = [("volume", D), ("centroid", Vec3),
                ("face", FaceGeometry * 6), ("cfl_length", D * 3)]


class EdgeGeometry(ctypes.Structure):
    _fields_ = [("length", D), ("normal", Vec3), ("tangent1", Vec3),
                ("tangent2", Vec3)]


class FluidFlux(ctypes.Structure):
    _fields_ = [("conserved", D * 5), ("conserved_jump", D * 5),
                ("velocity_jump", Vec3), ("normal_velocity", D * 2)]


class MaxwellFlux(ctypes.Structure):
    _fields_ = [("momentum", Vec3), ("alfven_diffusion_speed", D),
                ("magnetic_pressure_sum", D)]


class WindState(ctypes.Structure):
    _fields_ = [("density", D), ("pressure", D), ("velocity", D * 3),
                ("magnetic", D * 3)]


class WindSeries(ctypes.Structure):
    _fields_ = [("count", ctypes.c_size_t), ("time", PD),
                ("state", ctypes.POINTER(WindState)), ("reference", D * 3),
                ("time_offset", D), ("by_coefficient", D),
                ("bz_coefficient", D), ("bx_offset", D),
                ("enforce_bx_relation", I), ("linear_interpolation", I)]


# name -> (reference symbol, oracle symbol, restype, argtypes)
FUNCTIONS = {
    "up7_pdm": ("gamera_no_reconstruct_up7_pdm", "orc_up7_pdm", I,
                [PD, PD, D, PD, PD]),
    "cons_to_prim": ("gamera_no_conserved_to_primitive", "orc_cons_to_prim", I,
                     [PD, D, D, D, ctypes.POINTER(Primitive)]),
    "predict_cell": ("gamera_no_predict_cell", "orc_predict_cell", I,
                     [PD, PD, D, D, D, D, PD]),
    "fluid_flux": ("gamera_no_kinetic_fluid_flux", "orc_fluid_flux_kinetic", I,
                   [ctypes.POINTER(Primitive), D,
                    ctypes.POINTER(FaceGeometry), ctypes.POINTER(FluidFlux)]),
    "maxwell_flux": ("gamera_no_kinetic_maxwell_flux",
                     "orc_maxwell_flux_kinetic", I,
                     [ctypes.POINTER(Primitive), ctypes.POINTER(Vec3), D,
                      ctypes.POINTER(FaceGeometry), B, Vec3, D, B, D, PD,
                      ctypes.POINTER(MaxwellFlux)]),
    "hogs": ("gamera_no_apply_hogs", "orc_hogs", I,
             [ctypes.POINTER(FluidFlux), ctypes.POINTER(MaxwellFlux), D, D, B,
              D]),
    "emergency": ("gamera_no_apply_emergency_interface_diffusion",
                  "orc_emergency_diffusion", I,
                  [ctypes.POINTER(FluidFlux), ctypes.POINTER(Primitive), PD, PD,
                   B, D, ctypes.POINTER(B)]),
    "boris": ("gamera_no_apply_boris", "orc_apply_boris", I,
              [PD, PD, Vec3, Vec3, PD, Vec3, D, D, D, D, D]),
    "edge_emf": ("gamera_no_compute_edge_emf", "orc_edge_emf_formula", I,
                 [D, D, D, D, D, D, ctypes.POINTER(EdgeGeometry), D, B, D, D,
                  D, PD, PD]),
    "edge_solve": ("gamera_no_solve_edge_field", "orc_edge_field_solve", I,
                   [PD, PD, D, D, PD]),
    "normal_at_edge": ("gamera_no_interpolate_face_normal_to_edge",
                       "orc_face_normal_at_edge", I,
                       [PD, ctypes.POINTER(Vec3), ctypes.POINTER(EdgeGeometry),
                        PD]),
    "recover": ("gamera_no_flux_to_cell_field", None, I,
                [ctypes.POINTER(CellGeometry), PD, ctypes.POINTER(Vec3), PD]),
    "wind_sample": ("gamera_solar_wind_sample_at", "orc_wind_sample", I,
                    [ctypes.POINTER(WindSeries), PD, D,
                     ctypes.POINTER(WindState), PD]),
    "wind_weight": ("gamera_solar_wind_weight", "orc_wind_weight", I,
                    [ctypes.POINTER(WindSeries), PD, D, PD]),
    "inverse": ("gamera_cs_inverse", "orc_cube_inverse", I,
                [Vec3, ctypes.POINTER(I), PD, PD]),
}


def bind(lib, which):
    """Return {name: function} for 'ref' or 'orc' symbols of FUNCTIONS."""
    out = {}
    for name, (ref_sym, orc_sym, restype, argtypes) in FUNCTIONS.items():
        sym = ref_sym if which == "ref" else orc_sym
        if sym is None:
            continue
        fn = getattr(lib, sym)
        fn.restype = restype
        fn.argtypes = argtypes
        out[name] = fn
    if which == "orc":
        # the oracle's sampler does not report the ballistic delay
        out["wind_sample"].argtypes = [ctypes.POINTER(WindSeries), PD, D,
                                       ctypes.POINTER(WindState)]
        fn = lib.orc_recover_cell
        fn.restype = I
        fn.argtypes = [ctypes.POINTER(CellGeometry), PD, ctypes.POINTER(Vec3)]
        out["recover"] = fn
        fn = lib.orc_cube_neighbor
        fn.restype = I
        fn.argtypes = [I, I, ctypes.POINTER(I), ctypes.POINTER(I),
                       ctypes.POINTER(I)]
        out["neighbor"] = fn
    return out
