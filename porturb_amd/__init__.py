"""porturb_amd -- MI355X-native portUrb WENO-FV dycore step and TKE LES closure.

Python access to libporturb.so through its C ABI (include/porturb.h).  The library is the
product; this module only loads it, mirrors its structs with ctypes and offers small helpers
(device buffers, case configs) used by the benchmark driver and the tests.  There is no CPU
fallback: every call goes to the HIP library, and loading fails loudly when it is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libporturb.so")

PU_MAX_TRACERS = 16
PU_OK, PU_ERR_ARG, PU_ERR_HIP, PU_ERR_STATE, PU_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
PU_BC_PERIODIC, PU_BC_PRECURSOR, PU_BC_SOLID_WALL = 0, 1, 2
PU_FP64, PU_FP32 = 0, 1
PU_CASE_BUBBLE, PU_CASE_ABL_URBAN, PU_CASE_WINDFARM = 1, 2, 3


class PuConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xlen", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double),
        ("ord", C.c_int), ("num_tracers", C.c_int), ("idWV", C.c_int), ("idTKE", C.c_int),
        ("bc_x", C.c_int), ("bc_y", C.c_int), ("bc_z", C.c_int),
        ("weno_all", C.c_int), ("enable_gravity", C.c_int),
        ("latitude", C.c_double), ("cfl", C.c_double), ("immersed_power", C.c_double),
        ("R_d", C.c_double), ("R_v", C.c_double), ("cp_d", C.c_double), ("gamma_d", C.c_double),
        ("C0", C.c_double), ("grav", C.c_double), ("p0", C.c_double),
        ("tracer_positive", C.c_ubyte * PU_MAX_TRACERS),
        ("tracer_adds_mass", C.c_ubyte * PU_MAX_TRACERS),
        ("tracer_diffuse", C.c_ubyte * PU_MAX_TRACERS),
        ("kinematic_viscosity", C.c_double), ("dns", C.c_int),
        ("les_total_mult", C.c_double), ("les_dissipation_mult", C.c_double),
        ("les_shear_prod_mult", C.c_double),
        ("precision", C.c_int),
    ]


class PuCouplerState(C.Structure):
    _fields_ = [
        ("density_dry", C.c_void_p), ("uvel", C.c_void_p), ("vvel", C.c_void_p),
        ("wvel", C.c_void_p), ("temp", C.c_void_p),
        ("tracers", C.c_void_p * PU_MAX_TRACERS),
        ("immersed_proportion", C.c_void_p),
    ]


_lib = None


def lib():
    """Load libporturb.so (built in-tree by porturb_amd.build).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libporturb.so is not built: run `python -m porturb_amd.build` "
                           "(or __graft_entry__.build()) first; there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    P, I, D, S = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    cfgp, csp = C.POINTER(PuConfig), C.POINTER(PuCouplerState)
    sig = {
        "pu_create": (I, [cfgp, P, C.POINTER(P)]),
        "pu_destroy": (I, [P]),
        "pu_last_error": (C.c_char_p, []),
        "pu_build_info": (C.c_char_p, []),
        "pu_dycore_init": (I, [P, csp]),
        "pu_dycore_compute_time_step": (D, [P]),
        "pu_dycore_time_step": (I, [P, csp, D]),
        "pu_dycore_tendencies": (I, [P, csp, P]),
        "pu_dycore_flat_census": (I, [P, csp, P]),
        "pu_dycore_hydrostatics": (I, [P] + [P] * 6),
        "pu_dycore_immersed_masks": (I, [P, P, P]),
        "pu_les_closure_apply": (I, [P, csp, D]),
        "pu_diagnostics": (I, [P, csp, P, P, P]),
        "pu_alloc_device": (I, [S, C.POINTER(P)]),
        "pu_free_device": (I, [P]),
        "pu_memcpy_h2d": (I, [P, P, S]),
        "pu_memcpy_d2h": (I, [P, P, S]),
        "pu_memset_device": (I, [P, I, S]),
        "pu_synchronize": (I, [P]),
        "pu_case_init": (I, [cfgp, I, C.c_ulonglong, csp, P]),
        "pu_timer_start": (I, [P]),
        "pu_timer_stop": (I, [P, C.POINTER(D)]),
        "pu_kernel_timing": (I, [P, I]),
        "pu_kernel_time": (I, [P, C.c_char_p, C.POINTER(D), C.POINTER(C.c_longlong)]),
        "pu_les_closure_init": (I, [P, csp]),
        "pu_les_closure_set_immersed_halos": (I, [P, P]),
        "pu_fill_device": (I, [P, D, S]),
        "pu_memcpy_d2d": (I, [P, P, S]),
        "pu_measure_fma_rate": (I, [I, C.POINTER(D)]),
        "pu_measure_copy_rate": (I, [S, C.POINTER(D)]),
        "pu_driver_create": (I, [I, I, I, I, D, I, I, C.c_ulonglong, I, C.c_char_p, C.POINTER(P)]),
        "pu_driver_step": (I, [P, I]),
        "pu_driver_set_forcing": (I, [P, P]),
        "pu_driver_write_output": (I, [P, C.c_char_p, I, C.c_char_p, I]),
        "pu_driver_restart": (I, [P, C.c_char_p]),
        "pu_driver_elapsed_time": (D, [P]),
        "pu_driver_set_turbines": (I, [P, C.c_char_p, I, I, I]),
        "pu_driver_set_dyn_cycle": (I, [P, I]),
        "pu_driver_turbines": (P, [P]),
        "pu_driver_state": (I, [P, csp, C.POINTER(I)]),
        "pu_driver_context": (P, [P]),
        "pu_driver_tracer_name": (I, [P, I, C.c_char_p, I]),
        "pu_driver_last_error": (C.c_char_p, []),
        "pu_driver_destroy": (I, [P]),
        "pu_mlp_create": (I, [C.POINTER(P)]),
        "pu_mlp_destroy": (I, [P]),
        "pu_mlp_set_params": (I, [P, P]),
        "pu_mlp_get_params": (I, [P, P]),
        "pu_mlp_set_normalisation": (I, [P, P, C.c_float, P, P]),
        "pu_mlp_get_normalisation": (I, [P, P, P, P, P]),
        "pu_mlp_save": (I, [P, C.c_char_p]),
        "pu_mlp_load": (I, [C.c_char_p, C.POINTER(P)]),
        "pu_mlp_pack_inputs": (I, [P, C.c_longlong, P]),
        "pu_mlp_forward": (I, [P, P, P, I, P]),
        "pu_mlp_loss_grad": (I, [P, P, P, P, I, P, C.POINTER(C.c_float)]),
        "pu_mlp_train_step": (I, [P, P, P, P, I, C.c_float, C.POINTER(C.c_float)]),
        "pu_mlp_adam_apply": (I, [P, P, C.c_float]),
        "pu_sgs_filter": (I, [P, csp, I, P, P, P]),
        "pu_sgs_features": (I, [P, I, I, I, P, P, C.c_longlong, P]),
        "pu_les_closure_set_model": (I, [P, P]),
        "pu_les_closure_sgs": (I, [P, csp, D, P, P]),
        "pu_surface_fluxes_init": (I, [P, P, P, P, P]),
        "pu_surface_fluxes_apply": (I, [P, csp, D]),
        "pu_sponge_layer": (I, [P, csp, D, D, D]),
        "pu_geostrophic_wind_forcing": (I, [P, csp, D, D, D, D]),
        "pu_level_mean_wind": (I, [P, csp, I, I, P, P]),
        "pu_level_mean_wind_noalloc": (I, [P, csp, I, I, P, P]),
        "pu_add_uniform_wind": (I, [P, csp, D, D]),
        "pu_nc_create": (I, [C.c_char_p, I, C.POINTER(P)]),
        "pu_nc_open": (I, [C.c_char_p, C.POINTER(P)]),
        "pu_nc_def_dim": (I, [P, C.c_char_p, C.c_longlong]),
        "pu_nc_def_var": (I, [P, C.c_char_p, I, I, C.POINTER(C.c_char_p)]),
        "pu_nc_redef": (I, [P]),
        "pu_nc_enddef": (I, [P]),
        "pu_nc_inq_var": (I, [P, C.c_char_p, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_longlong)]),
        "pu_nc_put_host": (I, [P, C.c_char_p, P, I]),
        "pu_nc_get_host": (I, [P, C.c_char_p, P, I]),
        "pu_nc_put_device": (I, [P, C.c_char_p, P, P]),
        "pu_nc_get_device": (I, [P, C.c_char_p, P, P]),
        "pu_nc_format": (I, [P, C.POINTER(I)]),
        "pu_nc_close": (I, [P]),
        "pu_turbines_create": (I, [P, P, I, P, P, P, P, C.POINTER(P)]),
        "pu_turbines_apply": (I, [P, csp, D, P, P]),
        "pu_turbines_status": (I, [P, I, P]),
        "pu_turbines_trace": (I, [P, I, I, P]),
        "pu_turbines_destroy": (I, [P]),
        "pu_dycore_output_fields": (I, [P, csp, P, P, P]),
        "pu_dycore_restart": (I, [P, csp, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != PU_OK:
        raise RuntimeError("libporturb error %d: %s" % (rc, lib().pu_last_error().decode()))


# ------------------------------------------------------------------------------------------
# configuration
# ------------------------------------------------------------------------------------------
def physical_constants(R_d=287., cp_d=1003., R_v=461., p0=1.e5, grav=9.81):
    """The constants custom_modules::sc_init sets (sc_init.h:31-51)."""
    cv_d = cp_d - R_d
    gamma = cp_d / cv_d
    kappa = R_d / cp_d
    C0 = (R_d * p0 ** (-kappa)) ** gamma
    return dict(R_d=R_d, cp_d=cp_d, R_v=R_v, p0=p0, grav=grav, gamma_d=gamma, C0=C0)


def make_config(nx, ny, nz, xlen, ylen, zlen, ord=5, tracers=None, bc_x=PU_BC_PERIODIC,
                bc_y=PU_BC_PERIODIC, bc_z=PU_BC_SOLID_WALL, precision=PU_FP64, **opts):
    """Config dict.  tracers: list of (name, positive, adds_mass, diffuse) in tracer order;
    water_vapor is required by the dycore (option idWV), TKE by the LES closure."""
    if tracers is None:
        tracers = [("water_vapor", True, True, True)]
    names = [t[0] for t in tracers]
    cfg = dict(nx=nx, ny=ny, nz=nz, xlen=xlen, ylen=ylen, zlen=zlen, ord=ord,
               tracers=list(tracers), num_tracers=len(tracers),
               idWV=names.index("water_vapor"),
               idTKE=names.index("TKE") if "TKE" in names else -1,
               bc_x=bc_x, bc_y=bc_y, bc_z=bc_z, weno_all=1, enable_gravity=1, latitude=0.,
               cfl=0.6, immersed_power=5., kinematic_viscosity=0., dns=0, les_total_mult=1.,
               les_dissipation_mult=1., les_shear_prod_mult=1., precision=precision)
    cfg.update(physical_constants())
    cfg.update(opts)
    return cfg


def to_pu_config(cfg):
    c = PuConfig()
    for name, _ in PuConfig._fields_:
        if name in ("tracer_positive", "tracer_adds_mass", "tracer_diffuse"):
            continue
        setattr(c, name, cfg[name])
    for i, (_, pos, mass, diff) in enumerate(cfg["tracers"]):
        c.tracer_positive[i] = int(pos)
        c.tracer_adds_mass[i] = int(mass)
        c.tracer_diffuse[i] = int(diff)
    return c


# ------------------------------------------------------------------------------------------
# device buffers and the coupler state
# ------------------------------------------------------------------------------------------
class DeviceBuffer:
    """A device allocation made through the library (pu_alloc_device)."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(lib().pu_alloc_device(nbytes, C.byref(p)))
        self.ptr, self.nbytes = p.value, nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes == self.nbytes
        check(lib().pu_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        check(lib().pu_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().pu_free_device(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CouplerState:
    """Device-resident coupler entries density_dry, uvel, vvel, wvel, temp, tracers,
    immersed_proportion ([nz][ny][nx] float64 each), as core::DataManager holds them."""

    FIELDS = ("density_dry", "uvel", "vvel", "wvel", "temp")

    def __init__(self, cfg):
        self.cfg = cfg
        self.shape = (cfg["nz"], cfg["ny"], cfg["nx"])
        n = int(np.prod(self.shape)) * 8
        self.bufs = {f: DeviceBuffer(n) for f in self.FIELDS}
        self.tracers = [DeviceBuffer(n) for _ in range(cfg["num_tracers"])]
        self.immersed = DeviceBuffer(n)
        self.c = PuCouplerState()
        for f in self.FIELDS:
            setattr(self.c, f, self.bufs[f].ptr)
        for i, b in enumerate(self.tracers):
            self.c.tracers[i] = b.ptr
        self.c.immersed_proportion = self.immersed.ptr

    def upload(self, host):
        """host: array [5+ntr][nz][ny][nx] (coupler order) and host['imm'] optional."""
        state, imm = host
        for i, f in enumerate(self.FIELDS):
            self.bufs[f].upload(state[i])
        for t, b in enumerate(self.tracers):
            b.upload(state[5 + t])
        if imm is not None:
            self.immersed.upload(imm)

    def download(self):
        nv = 5 + len(self.tracers)
        out = np.empty((nv,) + self.shape)
        for i, f in enumerate(self.FIELDS):
            out[i] = self.bufs[f].download(np.float64, self.shape)
        for t, b in enumerate(self.tracers):
            out[5 + t] = b.download(np.float64, self.shape)
        return out

    def download_immersed(self):
        return self.immersed.download(np.float64, self.shape)


class Dycore:
    """A libporturb context: Dynamics_Euler_Stratified_WenoFV + LES_Closure for one config."""

    def __init__(self, cfg, stream=None):
        self.cfg = cfg
        self.pc = to_pu_config(cfg)
        p = C.c_void_p()
        check(lib().pu_create(C.byref(self.pc), stream, C.byref(p)))
        self.ctx = p.value

    def init(self, cs):
        check(lib().pu_dycore_init(self.ctx, C.byref(cs.c)))

    def compute_time_step(self):
        return lib().pu_dycore_compute_time_step(self.ctx)

    def time_step(self, cs, dt):
        check(lib().pu_dycore_time_step(self.ctx, C.byref(cs.c), dt))

    def les_apply(self, cs, dt):
        check(lib().pu_les_closure_apply(self.ctx, C.byref(cs.c), dt))

    def synchronize(self):
        check(lib().pu_synchronize(self.ctx))

    def tendencies(self, cs):
        nv = 5 + self.cfg["num_tracers"]
        shape = (nv, self.cfg["nz"], self.cfg["ny"], self.cfg["nx"])
        buf = DeviceBuffer(int(np.prod(shape)) * 8)
        check(lib().pu_dycore_tendencies(self.ctx, C.byref(cs.c), buf.ptr))
        out = buf.download(np.float64, shape)
        buf.free()
        return out

    def flat_census(self, cs):
        """Reconstruction branches of the first stage of a time step from `cs`: int64 [3][3], rows
        x, y, z, columns [limiter, flat stencil, linear]."""
        out = np.zeros(9, dtype=np.int64)
        check(lib().pu_dycore_flat_census(self.ctx, C.byref(cs.c), out.ctypes.data))
        return out.reshape(3, 3)

    def hydrostatics(self):
        hs = (self.cfg["ord"] - 1) // 2
        nzh, nze = self.cfg["nz"] + 2 * hs, self.cfg["nz"] + 1
        arrs = [np.zeros(nzh) for _ in range(3)] + [np.zeros(nze) for _ in range(3)]
        check(lib().pu_dycore_hydrostatics(self.ctx, *[a.ctypes.data for a in arrs]))
        return arrs

    def immersed_masks(self):
        shape = (self.cfg["nz"], self.cfg["ny"], self.cfg["nx"])
        any10 = np.zeros(shape, np.uint8)
        tau = np.zeros(shape, np.int32)
        check(lib().pu_dycore_immersed_masks(self.ctx, any10.ctypes.data, tau.ctypes.data))
        return any10, tau

    def surface_init(self, rough, temp, sfc_rough, sfc_temp):
        """Host arrays immersed_roughness, immersed_temp [nz][ny][nx], surface_roughness,
        surface_temp [ny][nx]; uploaded and turned into the halo'd boxes."""
        bufs = []
        for a in (rough, temp, sfc_rough, sfc_temp):
            a = np.ascontiguousarray(a, dtype=np.float64)
            b = DeviceBuffer(a.nbytes)
            b.upload(a)
            bufs.append(b)
        check(lib().pu_surface_fluxes_init(self.ctx, *[b.ptr for b in bufs]))
        for b in bufs:
            b.free()

    def surface_apply(self, cs, dt):
        check(lib().pu_surface_fluxes_apply(self.ctx, C.byref(cs.c), dt))

    def sponge_layer(self, cs, dt, time_scale, top_prop):
        check(lib().pu_sponge_layer(self.ctx, C.byref(cs.c), dt, time_scale, top_prop))

    def geostrophic_forcing(self, cs, dt, lat_g, u_g, v_g):
        check(lib().pu_geostrophic_wind_forcing(self.ctx, C.byref(cs.c), dt, lat_g, u_g, v_g))

    def level_mean_wind(self, cs, k_begin, k_end):
        """Horizontal means of uvel and vvel on levels k_begin .. k_end-1 (host arrays)."""
        n = max(k_end - k_begin, 0)
        u, v = np.zeros(n), np.zeros(n)
        check(lib().pu_level_mean_wind(self.ctx, C.byref(cs.c), k_begin, k_end, u.ctypes.data, v.ctypes.data))
        return u, v

    def level_mean_wind_noalloc(self, cs, k_begin, k_end, u_dev, v_dev):
        """pu_level_mean_wind_noalloc: the level means into device pointers u_dev, v_dev
        (stream-ordered; the library allocates nothing)."""
        check(lib().pu_level_mean_wind_noalloc(self.ctx, C.byref(cs.c), k_begin, k_end, u_dev, v_dev))

    def add_uniform_wind(self, cs, du, dv):
        check(lib().pu_add_uniform_wind(self.ctx, C.byref(cs.c), du, dv))

    def set_sgs_model(self, mlp):
        check(lib().pu_les_closure_set_model(self.ctx, mlp.h if mlp is not None else None))

    def les_init(self, cs):
        check(lib().pu_les_closure_init(self.ctx, C.byref(cs.c)))

    def set_immersed_halos(self, halos):
        """pu_les_closure_set_immersed_halos with a host array [nz+2][ny+2][nx+2] (uploaded)."""
        halos = np.ascontiguousarray(halos, dtype=np.float64)
        b = DeviceBuffer(halos.nbytes)
        b.upload(halos)
        check(lib().pu_les_closure_set_immersed_halos(self.ctx, b.ptr))
        self.synchronize()
        b.free()

    def sgs(self, cs, dt):
        """(tau [9][nz][ny][nx] float32, flags [nz][ny][nx] uint8) of the attached network."""
        shape = (self.cfg["nz"], self.cfg["ny"], self.cfg["nx"])
        n3 = int(np.prod(shape))
        T, Fl = DeviceBuffer(4 * 9 * n3), DeviceBuffer(n3)
        check(lib().pu_les_closure_sgs(self.ctx, C.byref(cs.c), dt, T.ptr, Fl.ptr))
        tau = T.download(np.float32, (9,) + shape)
        fl = Fl.download(np.uint8, shape)
        T.free()
        Fl.free()
        return tau, fl

    def diagnostics(self, cs):
        m, w, a = C.c_double(), C.c_double(), C.c_double()
        check(lib().pu_diagnostics(self.ctx, C.byref(cs.c), C.byref(m), C.byref(w), C.byref(a)))
        return m.value, w.value, a.value

    def timer_start(self):
        check(lib().pu_timer_start(self.ctx))

    def timer_stop(self):
        ms = C.c_double()
        check(lib().pu_timer_stop(self.ctx, C.byref(ms)))
        return ms.value

    def kernel_timing(self, on):
        check(lib().pu_kernel_timing(self.ctx, int(on)))

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_longlong()
        check(lib().pu_kernel_time(self.ctx, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def destroy(self):
        if self.ctx:
            lib().pu_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def measure_ceilings(copy_bytes=2 << 30):
    """The device's vector FMA rates (TFLOP/s, FP64 and non-packed FP32) and the HBM rate of the
    plain streaming copy (GB/s read + written), pu_measure_fma_rate / pu_measure_copy_rate."""
    f64, f32, bw = C.c_double(), C.c_double(), C.c_double()
    check(lib().pu_measure_fma_rate(PU_FP64, C.byref(f64)))
    check(lib().pu_measure_fma_rate(PU_FP32, C.byref(f32)))
    check(lib().pu_measure_copy_rate(copy_bytes, C.byref(bw)))
    return {"fma_fp64_TFLOPs": f64.value, "fma_fp32_TFLOPs": f32.value, "copy_GBps": bw.value}


def case_init(cfg, case_id, seed, cs):
    pc = to_pu_config(cfg)
    check(lib().pu_case_init(C.byref(pc), case_id, seed, C.byref(cs.c), cs.immersed.ptr))


# ------------------------------------------------------------------------------------------
# learned subgrid closure
# ------------------------------------------------------------------------------------------
SGS_NIN, SGS_NH, SGS_NOUT, SGS_NINP = 108, 64, 9, 112
SGS_NPARAM = SGS_NH * SGS_NIN + SGS_NH + SGS_NH * SGS_NH + SGS_NH + SGS_NOUT * SGS_NH + SGS_NOUT


class Mlp:
    """A pu_mlp: the 108-64-64-9 network, its normalisation and Adam state, on the device."""

    def __init__(self, params=None, handle=None):
        if handle is None:
            p = C.c_void_p()
            check(lib().pu_mlp_create(C.byref(p)))
            handle = p.value
        self.h = handle
        if params is not None:
            self.set_params(params)

    @classmethod
    def load(cls, path):
        p = C.c_void_p()
        check(lib().pu_mlp_load(path.encode(), C.byref(p)))
        return cls(handle=p.value)

    def save(self, path):
        check(lib().pu_mlp_save(self.h, path.encode()))

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        assert a.size == SGS_NPARAM
        check(lib().pu_mlp_set_params(self.h, a.ctypes.data))

    def get_params(self):
        a = np.zeros(SGS_NPARAM, np.float32)
        check(lib().pu_mlp_get_params(self.h, a.ctypes.data))
        return a

    def set_normalisation(self, feat_scale, xmax, out_mean, out_scale):
        fs, om, os_ = (np.ascontiguousarray(x, dtype=np.float32) for x in (feat_scale, out_mean, out_scale))
        check(lib().pu_mlp_set_normalisation(self.h, fs.ctypes.data, float(xmax), om.ctypes.data, os_.ctypes.data))

    def get_normalisation(self):
        fs, om, os_ = np.zeros(SGS_NIN, np.float32), np.zeros(SGS_NOUT, np.float32), np.zeros(SGS_NOUT, np.float32)
        xm = C.c_float()
        check(lib().pu_mlp_get_normalisation(self.h, fs.ctypes.data, C.byref(xm), om.ctypes.data, os_.ctypes.data))
        return fs, xm.value, om, os_

    def forward(self, Xbf, n, idx=None):
        """Xbf: DeviceBuffer [N][112] bf16; returns host [n][9]."""
        Y = DeviceBuffer(4 * SGS_NOUT * n)
        check(lib().pu_mlp_forward(self.h, Xbf.ptr, idx.ptr if idx is not None else None, n, Y.ptr))
        out = Y.download(np.float32, (n, SGS_NOUT))
        Y.free()
        return out

    def loss_grad(self, Xbf, T, n, idx=None):
        G = DeviceBuffer(4 * SGS_NPARAM)
        loss = C.c_float()
        check(lib().pu_mlp_loss_grad(self.h, Xbf.ptr, T.ptr, idx.ptr if idx is not None else None, n, G.ptr,
                                     C.byref(loss)))
        g = G.download(np.float32, (SGS_NPARAM,))
        G.free()
        return loss.value, g

    def train_step(self, Xbf, T, idx, n, lr, want_loss=False):
        loss = C.c_float()
        check(lib().pu_mlp_train_step(self.h, Xbf.ptr, T.ptr, idx.ptr if idx is not None else None, n, lr,
                                      C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def adam_apply(self, grad, lr):
        G = DeviceBuffer(4 * SGS_NPARAM)
        G.upload(np.ascontiguousarray(grad, dtype=np.float32))
        check(lib().pu_mlp_adam_apply(self.h, G.ptr, lr))
        G.free()

    def destroy(self):
        if self.h:
            lib().pu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def pack_inputs(X):
    """Host float [n][108] -> device bf16 [n][112] (DeviceBuffer)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n = X.shape[0]
    dX = DeviceBuffer(X.nbytes)
    dX.upload(X)
    out = DeviceBuffer(2 * SGS_NINP * n)
    check(lib().pu_mlp_pack_inputs(dX.ptr, n, out.ptr))
    dX.free()
    return out


def sgs_filter(dycore_ctx, cs_struct, cfg, r):
    """Box filter of a fine state: device buffers (les, imm, targets) of the coarse grid."""
    nxc, nyc, nzc = cfg["nx"] // r, cfg["ny"] // r, cfg["nz"] // r
    nh = (nxc + 2) * (nyc + 2) * (nzc + 2)
    les, imm, tg = DeviceBuffer(8 * 6 * nh), DeviceBuffer(8 * nh), DeviceBuffer(4 * 9 * nxc * nyc * nzc)
    check(lib().pu_sgs_filter(dycore_ctx, C.byref(cs_struct), r, les.ptr, imm.ptr, tg.ptr))
    return les, imm, tg


def sgs_features(les, nx, ny, nz, fmul, cells):
    """Inputs of the listed cells of a halo-1 box (DeviceBuffer) -> DeviceBuffer bf16 [n][112]."""
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    fm = np.ascontiguousarray(fmul, dtype=np.float32)
    dc = DeviceBuffer(cells.nbytes)
    dc.upload(cells)
    out = DeviceBuffer(2 * SGS_NINP * cells.size)
    check(lib().pu_sgs_features(les.ptr, nx, ny, nz, fm.ctypes.data, dc.ptr, cells.size, out.ptr))
    dc.free()
    return out


def bf16_to_float(a):
    """uint16/int16 bfloat16 bit patterns -> float32."""
    return (np.asarray(a).astype(np.uint16).astype(np.uint32) << 16).view(np.float32)


PU_NC_BYTE, PU_NC_CHAR, PU_NC_SHORT, PU_NC_INT, PU_NC_FLOAT, PU_NC_DOUBLE, PU_NC_UBYTE, PU_NC_INT64 = 1, 2, 3, 4, 5, 6, 7, 10
PU_NC_AUTO, PU_NC_CDF2, PU_NC_CDF5 = 0, 2, 5
_NC_HOST = {np.dtype(np.float64): PU_NC_DOUBLE, np.dtype(np.float32): PU_NC_FLOAT, np.dtype(np.int32): PU_NC_INT}


class NcFile:
    """NetCDF classic file through libporturb's reader/writer (pu_nc_*)."""

    def __init__(self, path, mode="r", format=PU_NC_AUTO):
        p = C.c_void_p()
        if mode == "w":
            check(lib().pu_nc_create(path.encode(), format, C.byref(p)))
        else:
            check(lib().pu_nc_open(path.encode(), C.byref(p)))
        self.f = p.value

    def def_dim(self, name, n):
        check(lib().pu_nc_def_dim(self.f, name.encode(), n))

    def def_var(self, name, nc_type, dims):
        arr = (C.c_char_p * max(len(dims), 1))(*[d.encode() for d in dims])
        check(lib().pu_nc_def_var(self.f, name.encode(), nc_type, len(dims), arr))

    def redef(self):
        check(lib().pu_nc_redef(self.f))

    def enddef(self):
        check(lib().pu_nc_enddef(self.f))

    def inq(self, name):
        t, nd, dims = C.c_int(), C.c_int(), (C.c_longlong * 8)()
        check(lib().pu_nc_inq_var(self.f, name.encode(), C.byref(t), C.byref(nd), dims))
        return t.value, tuple(dims[i] for i in range(nd.value))

    def put(self, name, arr):
        a = np.ascontiguousarray(arr)
        check(lib().pu_nc_put_host(self.f, name.encode(), a.ctypes.data, _NC_HOST[a.dtype]))

    def get(self, name, dtype=np.float64):
        _, shape = self.inq(name)
        a = np.empty(shape, dtype)
        check(lib().pu_nc_get_host(self.f, name.encode(), a.ctypes.data, _NC_HOST[np.dtype(dtype)]))
        return a

    def put_device(self, name, dptr, ctx):
        check(lib().pu_nc_put_device(self.f, name.encode(), dptr, ctx))

    def get_device(self, name, dptr, ctx):
        check(lib().pu_nc_get_device(self.f, name.encode(), dptr, ctx))

    def version(self):
        v = C.c_int()
        check(lib().pu_nc_format(self.f, C.byref(v)))
        return v.value

    def close(self):
        if self.f:
            f, self.f = self.f, None
            check(lib().pu_nc_close(f))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class PuTurbineRef(C.Structure):
    _fields_ = [("ntable", C.c_int), ("velocity_magnitude", C.c_void_p), ("thrust_coef", C.c_void_p),
                ("power_coef", C.c_void_p), ("power_megawatts", C.c_void_p), ("hub_height", C.c_double),
                ("blade_radius", C.c_double), ("max_yaw_speed", C.c_double), ("overhang", C.c_double),
                ("hub_radius", C.c_double), ("shaft_tilt", C.c_double)]


class PuTurbineOpts(C.Structure):
    _fields_ = [("initial_yaw", C.c_double), ("fixed_yaw", C.c_int), ("orig_C_T", C.c_int),
                ("have_upstream_dir", C.c_int), ("upstream_dir", C.c_double)]


class PuTurbineStatus(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("yaw", "u_samp_inertial", "v_samp_inertial", "power", "ct", "cp", "u_samp",
                                          "v_samp", "upstream_u", "upstream_v", "instant_mag0")] + [("nsteps", C.c_int)]


PU_TURB_POWER, PU_TURB_YAW, PU_TURB_U_SAMP, PU_TURB_V_SAMP, PU_TURB_CP, PU_TURB_CT = range(6)


class Turbines:
    """Actuator-disk turbines (pu_turbines_*) on a Dycore context."""

    def __init__(self, dycore, velmag, thrust_coef, power_coef, power_mw, hub_height, blade_radius, x, y,
                 max_yaw_speed=0.5 / 180 * np.pi, overhang=None, hub_radius=None, shaft_tilt=0.0, apply_thrust=None,
                 initial_yaw=0.0, fixed_yaw=False, orig_C_T=False, upstream_dir=None):
        self.tab = [np.ascontiguousarray(a, dtype=np.float64) for a in (velmag, thrust_coef, power_coef, power_mw)]
        ref = PuTurbineRef(len(self.tab[0]), *[a.ctypes.data for a in self.tab], hub_height, blade_radius,
                           max_yaw_speed, -0.1 * blade_radius if overhang is None else overhang,
                           0.03 * blade_radius if hub_radius is None else hub_radius, shaft_tilt)
        opts = PuTurbineOpts(initial_yaw, int(fixed_yaw), int(orig_C_T), int(upstream_dir is not None), upstream_dir or 0.0)
        xs, ys = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64)
        at = None if apply_thrust is None else np.ascontiguousarray(apply_thrust, np.uint8)
        p = C.c_void_p()
        check(lib().pu_turbines_create(dycore.ctx, C.byref(ref), len(xs), xs.ctypes.data, ys.ctypes.data,
                                       None if at is None else at.ctypes.data, C.byref(opts), C.byref(p)))
        self.t, self.n = p.value, len(xs)

    def apply(self, cs, dt, proj=None, samp=None):
        check(lib().pu_turbines_apply(self.t, C.byref(cs.c), dt, proj.ptr if proj else None, samp.ptr if samp else None))

    def status(self, i):
        st = PuTurbineStatus()
        check(lib().pu_turbines_status(self.t, i, C.byref(st)))
        return st

    def trace(self, i, which):
        n = self.status(i).nsteps
        out = np.zeros(n)
        check(lib().pu_turbines_trace(self.t, i, which, out.ctypes.data))
        return out

    def destroy(self):
        if self.t:
            lib().pu_turbines_destroy(self.t)
            self.t = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PuDriverForcing(C.Structure):
    _fields_ = [("geostrophic", C.c_int), ("lat_g", C.c_double), ("u_g", C.c_double), ("v_g", C.c_double),
                ("sponge", C.c_int), ("sponge_dt_multiple", C.c_double), ("sponge_top_prop", C.c_double),
                ("pg", C.c_int), ("pg_height", C.c_double), ("pg_u0", C.c_double), ("pg_v0", C.c_double),
                ("pg_tau", C.c_double), ("pg_update_cycles", C.c_ulong)]


class Driver:
    """The C++ benchmark driver (core::Coupler + Dynamics_Euler_Stratified_WenoFV + LES_Closure)
    for one case; the coupler's DataManager owns the state."""

    def __init__(self, case_id, nx, ny, nz, dx, ord=9, precision=PU_FP64, seed=1234, closure=True, sgs_model=None):
        p = C.c_void_p()
        rc = lib().pu_driver_create(case_id, nx, ny, nz, dx, ord, precision, seed, int(closure),
                                    sgs_model.encode() if sgs_model else None, C.byref(p))
        if rc != PU_OK:
            raise RuntimeError("pu_driver_create: %s %s" % (lib().pu_driver_last_error().decode(),
                                                          lib().pu_last_error().decode()))
        self.d = p.value
        self.shape = (nz, ny, nx)
        self.ctx = lib().pu_driver_context(self.d)

    def set_forcing(self, **kw):
        """Column forcing of the step: geostrophic (lat_g, u_g, v_g), sponge (sponge_dt_multiple,
        sponge_top_prop), pg = 1 (pg_tau, pg_update_cycles) or 2 (pg_height, pg_u0, pg_v0, pg_tau)."""
        f = PuDriverForcing(**kw)
        if lib().pu_driver_set_forcing(self.d, C.byref(f)) != PU_OK:
            raise ValueError("pu_driver_set_forcing: " + lib().pu_driver_last_error().decode())

    def write_output(self, prefix, double_precision=True):
        """Coupler::write_output_file; returns the file name."""
        buf = C.create_string_buffer(4096)
        if lib().pu_driver_write_output(self.d, prefix.encode(), int(double_precision), buf, 4096) != PU_OK:
            raise RuntimeError("pu_driver_write_output: " + lib().pu_driver_last_error().decode())
        return buf.value.decode()

    def restart(self, path):
        """Coupler::overwrite_with_restart from `path`."""
        if lib().pu_driver_restart(self.d, path.encode()) != PU_OK:
            raise RuntimeError("pu_driver_restart: " + lib().pu_driver_last_error().decode())

    def elapsed_time(self):
        return lib().pu_driver_elapsed_time(self.d)

    def set_turbines(self, turbine_file, ncols_x, nrows_y, orig_C_T=False):
        """modules::WindmillActuators (actuator disks) in the step: a ncols_x x nrows_y farm."""
        if lib().pu_driver_set_turbines(self.d, turbine_file.encode(), ncols_x, nrows_y, int(orig_C_T)) != PU_OK:
            raise RuntimeError("pu_driver_set_turbines: " + lib().pu_driver_last_error().decode())

    def turbine_status(self, i):
        st = PuTurbineStatus()
        check(lib().pu_turbines_status(lib().pu_driver_turbines(self.d), i, C.byref(st)))
        return st

    def set_dyn_cycle(self, n):
        """Physics step = n dycore steps (the reference ABL drivers' dyn_cycle)."""
        if lib().pu_driver_set_dyn_cycle(self.d, int(n)) != PU_OK:
            raise ValueError("pu_driver_set_dyn_cycle: " + lib().pu_driver_last_error().decode())

    def step(self, n=1):
        if lib().pu_driver_step(self.d, n) != PU_OK:
            raise RuntimeError("pu_driver_step: " + lib().pu_driver_last_error().decode())

    def tracer_names(self):
        cs, n = PuCouplerState(), C.c_int()
        check(lib().pu_driver_state(self.d, C.byref(cs), C.byref(n)))
        out = []
        for t in range(n.value):
            buf = C.create_string_buffer(64)
            check(lib().pu_driver_tracer_name(self.d, t, buf, 64))
            out.append(buf.value.decode())
        return out

    def download(self):
        """{name: array} of density_dry, uvel, vvel, wvel, temp and every tracer."""
        cs, n = PuCouplerState(), C.c_int()
        check(lib().pu_driver_state(self.d, C.byref(cs), C.byref(n)))
        out = {}
        nb = int(np.prod(self.shape)) * 8
        for f in CouplerState.FIELDS:
            a = np.empty(self.shape)
            check(lib().pu_memcpy_d2h(a.ctypes.data, getattr(cs, f), nb))
            out[f] = a
        for t, name in enumerate(self.tracer_names()):
            a = np.empty(self.shape)
            check(lib().pu_memcpy_d2h(a.ctypes.data, cs.tracers[t], nb))
            out[name] = a
        return out

    def field_pointers(self):
        """{name: device pointer} of density_dry, uvel, vvel, wvel, temp and every tracer."""
        cs, n = PuCouplerState(), C.c_int()
        check(lib().pu_driver_state(self.d, C.byref(cs), C.byref(n)))
        out = {f: getattr(cs, f) for f in CouplerState.FIELDS}
        for t, name in enumerate(self.tracer_names()):
            out[name] = cs.tracers[t]
        return out

    def download_field(self, name):
        a = np.empty(self.shape)
        check(lib().pu_memcpy_d2h(a.ctypes.data, self.field_pointers()[name], a.nbytes))
        return a

    def roll_y(self, shift):
        """Periodic shift of every coupler field by `shift` cells along y, on the device
        (np.roll(field, shift, axis=1) for [nz][ny][nx])."""
        nz, ny, nx = self.shape
        shift %= ny
        row = nx * 8
        tmp = DeviceBuffer(nz * ny * row)
        for ptr in self.field_pointers().values():
            for k in range(nz):
                src, dst = ptr + k * ny * row, tmp.ptr + k * ny * row
                check(lib().pu_memcpy_d2d(C.c_void_p(dst + shift * row), C.c_void_p(src), (ny - shift) * row))
                if shift:
                    check(lib().pu_memcpy_d2d(C.c_void_p(dst), C.c_void_p(src + (ny - shift) * row), shift * row))
            check(lib().pu_memcpy_d2d(C.c_void_p(ptr), C.c_void_p(tmp.ptr), nz * ny * row))
        check(lib().pu_synchronize(self.ctx))
        tmp.free()

    def synchronize(self):
        check(lib().pu_synchronize(self.ctx))

    def flat_census(self):
        """Reconstruction branches of the first stage of the next step on the current state:
        int64 [3][3], rows x, y, z, columns [limiter, flat stencil, linear]."""
        cs, n = PuCouplerState(), C.c_int()
        check(lib().pu_driver_state(self.d, C.byref(cs), C.byref(n)))
        out = np.zeros(9, dtype=np.int64)
        check(lib().pu_dycore_flat_census(self.ctx, C.byref(cs), out.ctypes.data))
        return out.reshape(3, 3)

    def timer_start(self):
        check(lib().pu_timer_start(self.ctx))

    def timer_stop(self):
        ms = C.c_double()
        check(lib().pu_timer_stop(self.ctx, C.byref(ms)))
        return ms.value

    def kernel_timing(self, on):
        check(lib().pu_kernel_timing(self.ctx, int(on)))

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_longlong()
        check(lib().pu_kernel_time(self.ctx, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def destroy(self):
        if self.d:
            lib().pu_driver_destroy(self.d)
            self.d = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
