"""ctypes access to the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the checker.
Configurations are the same dicts porturb_amd.make_config builds.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PORTURB_ORACLE_LIB selects another build of the same sources (the planted-mistake checks of
# tests/test_oracle_mutants.py load mutated copies this way)
LIB_PATH = os.environ.get("PORTURB_ORACLE_LIB", os.path.join(HERE, "liboracle.so"))
MAXT = 16


class OcConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xlen", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double),
        ("ord", C.c_int), ("num_tracers", C.c_int), ("idWV", C.c_int), ("idTKE", C.c_int),
        ("bc_x", C.c_int), ("bc_y", C.c_int), ("bc_z", C.c_int),
        ("weno_all", C.c_int), ("enable_gravity", C.c_int),
        ("latitude", C.c_double), ("cfl", C.c_double), ("immersed_power", C.c_double),
        ("R_d", C.c_double), ("R_v", C.c_double), ("cp_d", C.c_double), ("gamma_d", C.c_double),
        ("C0", C.c_double), ("grav", C.c_double), ("p0", C.c_double),
        ("tracer_positive", C.c_ubyte * MAXT), ("tracer_adds_mass", C.c_ubyte * MAXT),
        ("tracer_diffuse", C.c_ubyte * MAXT),
        ("kinematic_viscosity", C.c_double), ("dns", C.c_int),
        ("les_total_mult", C.c_double), ("les_dissipation_mult", C.c_double),
        ("les_shear_prod_mult", C.c_double),
        ("nthreads", C.c_int),
    ]


class OcSgsModel(C.Structure):
    _fields_ = [("params", C.c_void_p), ("feat_scale", C.c_void_p), ("xmax", C.c_float),
                ("out_mean", C.c_void_p), ("out_scale", C.c_void_p)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        cp = C.POINTER(OcConfig)
        L.oc_weno_tables.argtypes = [I] + [P] * 8
        L.oc_weno_tables.restype = I
        L.oc_weno_limited_edges.argtypes = [I, P, I, I, I, P, P]
        L.oc_weno_limited_edges.restype = I
        L.oc_weno_face_weights.argtypes = [I, P, I, I, I, P, P]
        L.oc_weno_face_weights.restype = I
        L.oc_weno_map_rs.argtypes = [D, D]
        L.oc_weno_map_rs.restype = D
        L.oc_dycore_compute_time_step.argtypes = [cp]
        L.oc_dycore_compute_time_step.restype = D
        L.oc_dycore_create.argtypes = [cp, P, P]
        L.oc_dycore_create.restype = P
        L.oc_dycore_destroy.argtypes = [P]
        L.oc_dycore_time_step.argtypes = [P, P, D]
        L.oc_dycore_time_step.restype = I
        L.oc_dycore_tendencies.argtypes = [P, P, P]
        L.oc_dycore_tendencies.restype = I
        L.oc_dycore_hydrostatics.argtypes = [P] * 7
        L.oc_dycore_hydrostatics.restype = I
        L.oc_dycore_immersed_masks.argtypes = [P, P, P]
        L.oc_dycore_immersed_masks.restype = I
        L.oc_dycore_output_fields.argtypes = [P, P, P, P, P]
        L.oc_dycore_output_fields.restype = I
        L.oc_les_apply.argtypes = [cp, P, P, D]
        L.oc_les_apply.restype = I
        L.oc_total_mass.argtypes = [cp, P]
        L.oc_total_mass.restype = D
        F, LG = C.c_float, C.c_long
        L.oc_bf16_round.argtypes = [F]
        L.oc_bf16_round.restype = F
        L.oc_relu.argtypes = [D]
        L.oc_relu.restype = D
        L.oc_relu_deriv.argtypes = [D]
        L.oc_relu_deriv.restype = D
        L.oc_mlp_forward.argtypes = [P, P, LG, I, P]
        L.oc_mlp_forward.restype = None
        L.oc_mlp_loss_grad.argtypes = [P, P, P, LG, I, P]
        L.oc_mlp_loss_grad.restype = D
        L.oc_mlp_adam.argtypes = [P, P, P, P, LG, I, F, F, F, F]
        L.oc_mlp_adam.restype = None
        L.oc_sgs_filter.argtypes = [cp, P, P, I, P, P, P]
        L.oc_sgs_filter.restype = I
        L.oc_sgs_fmul.argtypes = [P, D, P]
        L.oc_sgs_fmul.restype = None
        L.oc_sgs_features.argtypes = [P, I, I, I, P, P, LG, P]
        L.oc_sgs_features.restype = I
        L.oc_sgs_infer.argtypes = [C.POINTER(OcSgsModel), P, P, I, I, I, D, D, D, D, I, P, P]
        L.oc_sgs_infer.restype = None
        L.oc_les_apply_learned.argtypes = [cp, P, P, D, C.POINTER(OcSgsModel), I, P, P]
        L.oc_les_apply_learned.restype = I
        L.oc_surface_fluxes.argtypes = [cp, P, P, P, P, P, P, D]
        L.oc_surface_fluxes.restype = I
        L.oc_turbine_interp.argtypes = [P, P, I, D]
        L.oc_turbine_interp.restype = D
        L.oc_turbine_blade_weight.argtypes = [D, D]
        L.oc_turbine_blade_weight.restype = D
        L.oc_turbine_yaw_tend.argtypes = [C.POINTER(OcTurbine), D, D, D, D]
        L.oc_turbine_yaw_tend.restype = D
        L.oc_turbines_apply.argtypes = [cp, P, C.POINTER(OcTurbineRef), C.POINTER(OcTurbine), I,
                                        C.POINTER(OcTurbineOpts), D, P, P]
        L.oc_turbines_apply.restype = I
        L.oc_sponge_layer.argtypes = [cp, P, D, D, D]
        L.oc_sponge_layer.restype = I
        L.oc_geostrophic_forcing.argtypes = [cp, P, D, D, D, D]
        L.oc_geostrophic_forcing.restype = I
        L.oc_uniform_pg_forcing.argtypes = [cp, P, D, D, C.c_ulong, C.POINTER(OcPgState)]
        L.oc_uniform_pg_forcing.restype = I
        L.oc_uniform_pg_forcing_height.argtypes = [cp, P, D, D, D, D, D]
        L.oc_uniform_pg_forcing_height.restype = I
        _lib = L
    return _lib


def to_oc_config(cfg, nthreads=0):
    c = OcConfig()
    for name, _ in OcConfig._fields_:
        if name in ("tracer_positive", "tracer_adds_mass", "tracer_diffuse", "nthreads"):
            continue
        setattr(c, name, cfg[name])
    for i, (_, pos, mass, diff) in enumerate(cfg["tracers"]):
        c.tracer_positive[i] = int(pos)
        c.tracer_adds_mass[i] = int(mass)
        c.tracer_diffuse[i] = int(diff)
    c.nthreads = nthreads
    return c


def weno_tables(ord_):
    hs = (ord_ - 1) // 2
    N = S = hs + 1
    C_ = np.zeros((S, N, N)); fL = np.zeros((S, ord_)); fR = np.zeros((S, ord_))
    dL = np.zeros(S); dR = np.zeros(S); lL = np.zeros(ord_); lR = np.zeros(ord_); tv = np.zeros((N, N))
    rc = lib().oc_weno_tables(ord_, *[a.ctypes.data for a in (C_, fL, fR, dL, dR, lL, lR, tv)])
    assert rc == 0
    return dict(C=C_, faceL=fL, faceR=fR, idealL=dL, idealR=dR, linL=lL, linR=lR, tv=tv)


def weno_edges(ord_, stencil, do_map=True, imm_L=False, imm_R=False):
    s = np.ascontiguousarray(stencil, dtype=np.float64)
    qL, qR = C.c_double(), C.c_double()
    rc = lib().oc_weno_limited_edges(ord_, s.ctypes.data, int(do_map), int(imm_L), int(imm_R),
                                     C.byref(qL), C.byref(qR))
    assert rc == 0
    return qL.value, qR.value


def weno_face_weights(ord_, stencil, do_map=True, imm_L=False, imm_R=False):
    """Final normalised weights (left face, right face) of the limiter on one stencil."""
    s = np.ascontiguousarray(stencil, dtype=np.float64)
    S = (ord_ + 1) // 2
    wL, wR = np.zeros(S), np.zeros(S)
    rc = lib().oc_weno_face_weights(ord_, s.ctypes.data, int(do_map), int(imm_L), int(imm_R),
                                    wL.ctypes.data, wR.ctypes.data)
    assert rc == 0
    return wL, wR


def weno_map_rs(w, d):
    return lib().oc_weno_map_rs(float(w), float(d))


class Dycore:
    """The oracle's Dynamics_Euler_Stratified_WenoFV for one config and initial coupler state."""

    def __init__(self, cfg, coupler, imm, nthreads=0):
        self.cfg = cfg
        self.oc = to_oc_config(cfg, nthreads)
        self.coupler = np.ascontiguousarray(coupler, dtype=np.float64)
        self.imm = np.ascontiguousarray(imm, dtype=np.float64)
        self.h = lib().oc_dycore_create(C.byref(self.oc), self.coupler.ctypes.data, self.imm.ctypes.data)
        assert self.h, "oc_dycore_create failed"

    def time_step(self, coupler, dt):
        assert coupler.flags["C_CONTIGUOUS"] and coupler.dtype == np.float64
        assert lib().oc_dycore_time_step(self.h, coupler.ctypes.data, dt) == 0

    def tendencies(self, coupler):
        c = np.ascontiguousarray(coupler, dtype=np.float64)
        out = np.zeros_like(c)
        assert lib().oc_dycore_tendencies(self.h, c.ctypes.data, out.ctypes.data) == 0
        return out

    def hydrostatics(self):
        hs = (self.cfg["ord"] - 1) // 2
        nzh, nze = self.cfg["nz"] + 2 * hs, self.cfg["nz"] + 1
        arrs = [np.zeros(nzh) for _ in range(3)] + [np.zeros(nze) for _ in range(3)]
        lib().oc_dycore_hydrostatics(self.h, *[a.ctypes.data for a in arrs])
        return arrs

    def output_fields(self, coupler):
        """density_pert, theta_pert, pressure_pert of the dycore's output hook."""
        c = np.ascontiguousarray(coupler, dtype=np.float64)
        shape = (self.cfg["nz"], self.cfg["ny"], self.cfg["nx"])
        out = [np.zeros(shape) for _ in range(3)]
        assert lib().oc_dycore_output_fields(self.h, c.ctypes.data, *[a.ctypes.data for a in out]) == 0
        return out

    def immersed_masks(self):
        shape = (self.cfg["nz"], self.cfg["ny"], self.cfg["nx"])
        any10 = np.zeros(shape, np.uint8)
        tau = np.zeros(shape, np.int32)
        lib().oc_dycore_immersed_masks(self.h, any10.ctypes.data, tau.ctypes.data)
        return any10, tau

    def compute_time_step(self):
        return lib().oc_dycore_compute_time_step(C.byref(self.oc))

    def __del__(self):
        try:
            if self.h:
                lib().oc_dycore_destroy(self.h)
                self.h = None
        except Exception:
            pass


def les_apply(cfg, coupler, imm, dt, nthreads=0):
    oc = to_oc_config(cfg, nthreads)
    imm = np.ascontiguousarray(imm, dtype=np.float64)
    assert coupler.flags["C_CONTIGUOUS"] and coupler.dtype == np.float64
    assert lib().oc_les_apply(C.byref(oc), coupler.ctypes.data, imm.ctypes.data, dt) == 0


def total_mass(cfg, coupler):
    oc = to_oc_config(cfg)
    c = np.ascontiguousarray(coupler, dtype=np.float64)
    return lib().oc_total_mass(C.byref(oc), c.ctypes.data)


# --------------------------------------------------------------------------------------------
# learned subgrid closure (sgs.c)
# --------------------------------------------------------------------------------------------
NIN, NH, NOUT = 108, 64, 9
NPARAM = NH * NIN + NH + NH * NH + NH + NOUT * NH + NOUT


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def bf16_round(x):
    return lib().oc_bf16_round(float(x))


def mlp_forward(params, X, bf16=True):
    P, X = _f32(params), _f32(X)
    n = X.shape[0]
    Y = np.zeros((n, NOUT), np.float32)
    lib().oc_mlp_forward(P.ctypes.data, X.ctypes.data, n, int(bf16), Y.ctypes.data)
    return Y


def mlp_loss_grad(params, X, T, bf16=True):
    P, X, T = _f32(params), _f32(X), _f32(T)
    G = np.zeros(NPARAM)
    loss = lib().oc_mlp_loss_grad(P.ctypes.data, X.ctypes.data, T.ctypes.data, X.shape[0], int(bf16), G.ctypes.data)
    return loss, G


def mlp_adam(params, grad, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """In place on params, m, v (float32 arrays)."""
    g = _f32(grad)
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    lib().oc_mlp_adam(params.ctypes.data, g.ctypes.data, m.ctypes.data, v.ctypes.data, params.size, t,
                      lr, b1, b2, eps)


def sgs_filter(cfg, coupler, imm, r):
    oc = to_oc_config(cfg)
    nxc, nyc, nzc = cfg["nx"] // r, cfg["ny"] // r, cfg["nz"] // r
    les = np.zeros((6, nzc + 2, nyc + 2, nxc + 2))
    immc = np.zeros((nzc + 2, nyc + 2, nxc + 2))
    tg = np.zeros((9, nzc, nyc, nxc), np.float32)
    c = np.ascontiguousarray(coupler, dtype=np.float64)
    im = np.ascontiguousarray(imm, dtype=np.float64)
    assert lib().oc_sgs_filter(C.byref(oc), c.ctypes.data, im.ctypes.data, r, les.ctypes.data, immc.ctypes.data,
                               tg.ctypes.data) == 0
    return les, immc, tg


def sgs_fmul(feat_scale, delta):
    fs = _f32(feat_scale)
    out = np.zeros(NIN, np.float32)
    lib().oc_sgs_fmul(fs.ctypes.data, delta, out.ctypes.data)
    return out


def sgs_features(les, fmul, cells):
    les = np.ascontiguousarray(les, dtype=np.float64)
    nz, ny, nx = les.shape[1] - 2, les.shape[2] - 2, les.shape[3] - 2
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    fm = _f32(fmul)
    X = np.zeros((cells.size, NIN), np.float32)
    assert lib().oc_sgs_features(les.ctypes.data, nx, ny, nz, fm.ctypes.data, cells.ctypes.data, cells.size,
                                 X.ctypes.data) == 0
    return X


class SgsModel:
    """Parameters and normalisation of a learned closure, as the oracle reads them."""

    def __init__(self, params, feat_scale, xmax, out_mean, out_scale):
        self.params, self.feat_scale = _f32(params), _f32(feat_scale)
        self.out_mean, self.out_scale = _f32(out_mean), _f32(out_scale)
        self.xmax = float(xmax)
        self.s = OcSgsModel(self.params.ctypes.data, self.feat_scale.ctypes.data, self.xmax,
                            self.out_mean.ctypes.data, self.out_scale.ctypes.data)


def sgs_infer(model, les, imm, dx, dy, dz, nu_max, bf16=True):
    les = np.ascontiguousarray(les, dtype=np.float64)
    imm = np.ascontiguousarray(imm, dtype=np.float64)
    nz, ny, nx = les.shape[1] - 2, les.shape[2] - 2, les.shape[3] - 2
    tau = np.zeros((9, nz, ny, nx), np.float32)
    flags = np.zeros((nz, ny, nx), np.uint8)
    lib().oc_sgs_infer(C.byref(model.s), les.ctypes.data, imm.ctypes.data, nx, ny, nz, dx, dy, dz, nu_max,
                       int(bf16), tau.ctypes.data, flags.ctypes.data)
    return tau, flags


def les_apply_learned(cfg, coupler, imm, dt, model, bf16=True, nthreads=0):
    """LES closure step with the learned fluxes; returns (tau, flags) of the cells."""
    oc = to_oc_config(cfg, nthreads)
    imm = np.ascontiguousarray(imm, dtype=np.float64)
    assert coupler.flags["C_CONTIGUOUS"] and coupler.dtype == np.float64
    shape = (cfg["nz"], cfg["ny"], cfg["nx"])
    tau = np.zeros((9,) + shape, np.float32)
    flags = np.zeros(shape, np.uint8)
    assert lib().oc_les_apply_learned(C.byref(oc), coupler.ctypes.data, imm.ctypes.data, dt, C.byref(model.s),
                                      int(bf16), tau.ctypes.data, flags.ctypes.data) == 0
    return tau, flags


class OcTurbineRef(C.Structure):
    _fields_ = [("ntab", C.c_int), ("velmag", C.c_void_p), ("thrust_coef", C.c_void_p), ("power_coef", C.c_void_p),
                ("power_mw", C.c_void_p), ("hub_height", C.c_double), ("blade_radius", C.c_double),
                ("max_yaw_speed", C.c_double), ("overhang", C.c_double), ("hub_radius", C.c_double),
                ("shaft_tilt", C.c_double)]


class OcTurbine(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("base_x", "base_y", "yaw", "u_inertial", "v_inertial", "yaw_uavg",
                                          "yaw_vavg")] + [("apply_thrust", C.c_int)] + \
               [(n, C.c_double) for n in ("power", "ct", "cp", "u_samp", "v_samp", "upstream_u", "upstream_v")]


class OcTurbineOpts(C.Structure):
    _fields_ = [("fixed_yaw", C.c_int), ("orig_C_T", C.c_int), ("have_upstream_dir", C.c_int),
                ("upstream_dir", C.c_double)]


class TurbineRef:
    """A reference turbine: table arrays (kept alive here) plus geometry."""

    def __init__(self, velmag, thrust_coef, power_coef, power_mw, hub_height, blade_radius, max_yaw_speed=0.5 / 180 * np.pi,
                 overhang=None, hub_radius=None, shaft_tilt=0.0):
        self.arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in (velmag, thrust_coef, power_coef, power_mw)]
        self.s = OcTurbineRef(len(self.arrays[0]), *[a.ctypes.data for a in self.arrays], hub_height, blade_radius,
                              max_yaw_speed, -0.1 * blade_radius if overhang is None else overhang,
                              0.03 * blade_radius if hub_radius is None else hub_radius, shaft_tilt)


def turbines_apply(cfg, coupler, ref, turbs, dt, fixed_yaw=False, orig_C_T=False, upstream_dir=None):
    """modules::WindmillActuators::apply (actuator disks) in place; turbs is a ctypes array of
    OcTurbine (state carried between calls).  Returns (proj_weight_tot, samp_weight_tot)."""
    oc = to_oc_config(cfg)
    assert coupler.flags["C_CONTIGUOUS"] and coupler.dtype == np.float64
    shape = (cfg["nz"], cfg["ny"], cfg["nx"])
    pw, sw = np.zeros(shape), np.zeros(shape)
    o = OcTurbineOpts(int(fixed_yaw), int(orig_C_T), int(upstream_dir is not None), upstream_dir or 0.0)
    assert lib().oc_turbines_apply(C.byref(oc), coupler.ctypes.data, C.byref(ref.s), turbs, len(turbs), C.byref(o), dt,
                                   pw.ctypes.data, sw.ctypes.data) == 0
    return pw, sw


class OcPgState(C.Structure):
    """Options the uniform pressure-gradient forcing keeps between calls."""
    _fields_ = [("counter", C.c_ulong), ("u0", C.c_double), ("v0", C.c_double), ("u", C.c_double),
                ("v", C.c_double)]


def _coupler_call(fn, cfg, coupler, *args):
    oc = to_oc_config(cfg)
    assert coupler.flags["C_CONTIGUOUS"] and coupler.dtype == np.float64
    return fn(C.byref(oc), coupler.ctypes.data, *args)


def sponge_layer(cfg, coupler, dt, time_scale, top_prop):
    """modules::sponge_layer in place on the coupler state."""
    assert _coupler_call(lib().oc_sponge_layer, cfg, coupler, dt, time_scale, top_prop) == 0


def geostrophic_forcing(cfg, coupler, dt, lat_g, u_g, v_g):
    """modules::geostrophic_wind_forcing in place on the coupler state."""
    assert _coupler_call(lib().oc_geostrophic_forcing, cfg, coupler, dt, lat_g, u_g, v_g) == 0


def uniform_pg_forcing(cfg, coupler, dt, tau, update_cycles, state):
    """modules::uniform_pg_wind_forcing in place; `state` (OcPgState) carries the options."""
    assert _coupler_call(lib().oc_uniform_pg_forcing, cfg, coupler, dt, tau, update_cycles, C.byref(state)) == 0


def uniform_pg_forcing_height(cfg, coupler, dt, height, u0, v0, tau):
    """modules::uniform_pg_wind_forcing_height in place; returns False for a height the
    reference cannot serve (below the first cell centre)."""
    return _coupler_call(lib().oc_uniform_pg_forcing_height, cfg, coupler, dt, height, u0, v0, tau) == 0


def surface_fluxes(cfg, coupler, prop, rough, temp, sfc_rough, sfc_temp, dt):
    """modules::apply_surface_fluxes in place on the coupler state (float64 arrays)."""
    oc = to_oc_config(cfg)
    assert coupler.flags["C_CONTIGUOUS"] and coupler.dtype == np.float64
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (prop, rough, temp, sfc_rough, sfc_temp)]
    assert lib().oc_surface_fluxes(C.byref(oc), coupler.ctypes.data, *[a.ctypes.data for a in arrs], dt) == 0
