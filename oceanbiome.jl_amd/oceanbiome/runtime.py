"""ctypes binding of libobme_hip.so and the HIP-backed biogeochemistry.

`HIPBiogeochemistry(bgc, precision)` is the device counterpart of a
`Biogeochemistry` (NPZD / LOBSTER): it owns one library handle and exposes
the Oceananigans.Biogeochemistry hooks the library implements
(update_biogeochemical_state, update_tendencies) plus the fused stepper.
Tracer, tendency and PAR fields are torch tensors on the GPU holding
Oceananigans parent arrays (shape (Nz+2Hz, Ny+2Hy, Nx+2Hx)).

The library must be built (oceanbiome.jl_amd/csrc/build.sh); there is no
CPU fallback: a missing library or a missing GPU raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# OBME_LIBRARY names another build of the library (same ABI) to load instead
LIB_PATH = os.environ.get("OBME_LIBRARY") or os.path.join(HERE, "libobme_hip.so")

_lib = None


class OBMEError(RuntimeError):
    pass


def library():
    """Load libobme_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OBMEError(f"{LIB_PATH} not found: build it with oceanbiome.jl_amd/csrc/build.sh")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "obme_create": (C.c_int32, [C.POINTER(abi.Config), C.POINTER(vp)]),
        "obme_destroy": (C.c_int32, [vp]),
        "obme_last_error": (C.c_char_p, []),
        "obme_version": (C.c_char_p, []),
        "obme_required_tracers": (C.c_int32, [vp, C.c_char_p, C.c_int32]),
        "obme_update_state": (C.c_int32, [vp, C.POINTER(vp), vp, C.c_double, vp]),
        "obme_add_tendencies": (C.c_int32, [vp, C.POINTER(vp), C.POINTER(vp), vp, vp, vp, C.c_double, vp]),
        "obme_add_tendencies_with_sediment": (C.c_int32, [vp, C.POINTER(vp), C.POINTER(vp), vp, vp, vp, vp,
                                                          C.c_double, vp]),
        "obme_step": (C.c_int32, [vp, C.POINTER(vp), vp, vp, C.c_int32, C.c_double, C.c_double, vp]),
        "obme_step_with_sediment": (C.c_int32, [vp, C.POINTER(vp), vp, vp, vp, C.c_int32, C.c_double, C.c_double,
                                                vp]),
        "obme_step_noalloc": (C.c_int32, [vp, C.POINTER(vp), vp, vp, C.c_int32, C.c_double, C.c_double, vp]),
        "obme_sweep_configure": (C.c_int32, [vp, C.c_int32]),
        "obme_sweep_waves": (C.c_int32, [vp]),
        "obme_autotune": (C.c_int32, [vp, C.POINTER(vp), vp, vp, C.c_int32, C.c_double, vp]),
        "obme_sediment_update_state": (C.c_int32, [vp, C.POINTER(vp), vp, C.c_double, vp]),
        "obme_co2_flux": (C.c_int32, [vp, vp, vp, vp, vp, vp, C.c_double, vp, vp, C.c_int64, C.c_int32, vp]),
        "obme_abi_sizeof": (C.c_int64, [C.c_int32]),
        "obme_set_kernel_timing": (C.c_int32, [vp, C.c_int32]),
        "obme_batch_create": (C.c_int32, [C.POINTER(abi.Config), C.c_int32, C.POINTER(C.c_int32), C.c_double,
                                          C.POINTER(abi.BatchParams), C.c_double, C.POINTER(vp)]),
        "obme_batch_destroy": (C.c_int32, [vp]),
        "obme_batch_step": (C.c_int32, [vp, C.POINTER(vp), C.c_int32, C.c_double, C.c_double, vp]),
        "obme_batch_step_with_sediment": (C.c_int32, [vp, C.POINTER(vp), vp, C.c_int32, C.c_double, C.c_double,
                                                      vp]),
        "obme_batch_configure": (C.c_int32, [vp, C.c_int32, C.c_int32]),
        "obme_batch_autotune": (C.c_int32, [vp, C.POINTER(vp), C.c_int32, C.c_double, vp]),
        "obme_batch_describe": (C.c_int32, [vp, C.c_char_p, C.c_int32]),
        "obme_pisces_create": (C.c_int32, [C.POINTER(abi.PiscesConfig), C.POINTER(vp)]),
        "obme_pisces_destroy": (C.c_int32, [vp]),
        "obme_pisces_update_state": (C.c_int32, [vp, C.POINTER(vp), vp, C.POINTER(vp), vp, vp, vp, C.c_double,
                                                 vp]),
        "obme_pisces_add_tendencies": (C.c_int32, [vp, C.POINTER(vp), C.POINTER(vp), C.c_double, vp]),
        "obme_pisces_step": (C.c_int32, [vp, C.POINTER(vp), C.c_int32, C.c_double, C.c_double, vp]),
        "obme_pisces_step_with_sediment": (C.c_int32, [vp, C.POINTER(vp), vp, C.c_int32, C.c_double, C.c_double,
                                                       vp]),
        "obme_pisces_sediment_update_state": (C.c_int32, [vp, C.POINTER(vp), vp, C.c_double, vp]),
        "obme_kernel_timing": (C.c_int32, [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED_SYMBOLS = (
    "obme_create", "obme_destroy", "obme_last_error", "obme_version", "obme_required_tracers",
    "obme_update_state", "obme_add_tendencies", "obme_add_tendencies_with_sediment", "obme_step",
    "obme_step_with_sediment", "obme_step_noalloc",
    "obme_sweep_configure", "obme_sweep_waves", "obme_autotune",
    "obme_sediment_update_state", "obme_co2_flux", "obme_abi_sizeof",
    "obme_set_kernel_timing", "obme_kernel_timing", "obme_batch_create", "obme_batch_destroy", "obme_batch_step",
    "obme_batch_step_with_sediment",
    "obme_batch_configure", "obme_batch_autotune", "obme_batch_describe", "obme_pisces_create",
    "obme_pisces_destroy", "obme_pisces_update_state", "obme_pisces_add_tendencies", "obme_pisces_step",
    "obme_pisces_step_with_sediment", "obme_pisces_sediment_update_state",
)

KERNEL_SWEEP = 0      # OBME_KERNEL_SWEEP: the column sweep
KERNEL_SURFACE = 1    # OBME_KERNEL_SURFACE: the per-column CO2 surface kernel


def _check(status):
    if status != abi.OBME_OK:
        msg = library().obme_last_error().decode()
        raise OBMEError(f"obme status {status}: {msg}")


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _ptr_array(ts):
    return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class HIPBiogeochemistry:
    """Device biogeochemistry for a Biogeochemistry model description."""

    def __init__(self, bgc, precision="fp64", surface_array=None):
        import torch
        if not torch.cuda.is_available():
            raise OBMEError("HIPBiogeochemistry needs a GPU")
        self.bgc = bgc
        self.precision = abi.OBME_FP64 if precision == "fp64" else abi.OBME_FP32_MIXED
        self.dtype = torch.float64 if self.precision == abi.OBME_FP64 else torch.float32
        self.cfg = bgc.config(self.precision)
        self._surface = None
        if surface_array is not None:
            self._surface = surface_array.to(device="cuda", dtype=self.dtype).contiguous()
            self.cfg.par.surface_mode = abi.OBME_SURFACE_ARRAY
            self.cfg.par.surface_array = self._surface.data_ptr()
        h = C.c_void_p()
        _check(library().obme_create(C.byref(self.cfg), C.byref(h)))
        self.handle = h
        # the sediment's prognostic fields (BiogeochemicalSediment.fields):
        # InstantRemineralisation's storage (Ny, Nx), SimpleMultiG's Ns, Nf,
        # Nr (3, Ny, Nx)
        self.sediment_storage = None
        if getattr(bgc, "sediment", None) is not None:
            g = bgc.grid
            nf = len(bgc.sediment.required_sediment_fields())
            shape = (g.Ny, g.Nx) if nf == 1 else (nf, g.Ny, g.Nx)
            self.sediment_storage = torch.zeros(shape, dtype=torch.float64, device="cuda")

    def close(self):
        if getattr(self, "handle", None):
            library().obme_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- Oceananigans.Biogeochemistry API ---------------------------------------
    def required_biogeochemical_tracers(self):
        buf = C.create_string_buffer(512)
        n = library().obme_required_tracers(self.handle, buf, 512)
        if n < 0:
            raise OBMEError("obme_required_tracers failed")
        return tuple(buf.value.decode().split(";"))

    def _check_fields(self, fields):
        g = self.bgc.grid
        for f in fields:
            if f.dtype != self.dtype or tuple(f.shape) != g.parent_shape or not f.is_cuda or not f.is_contiguous():
                raise OBMEError(f"field must be a contiguous {self.dtype} CUDA tensor of shape {g.parent_shape}")

    def update_biogeochemical_state(self, tracers, PAR, t=0.0):
        """update_biogeochemical_state!: the negative-scaling modifier first
        (in place on the tracers, when configured), then the light step
        (Light/2band.jl:148-155): PAR interior from P."""
        self._check_fields(list(tracers) + [PAR])
        _check(library().obme_update_state(self.handle, _ptr_array(tracers), _ptr(PAR), t, _stream()))

    def update_tendencies(self, tracers, G, PAR=None, T=None, S=None, t=0.0):
        """Add BGC transitions, drift divergence, the top gas fluxes and the
        sediment's coupled tendencies into G (a SimpleMultiG sediment's from
        its fields, self.sediment_storage)."""
        self._check_fields(list(tracers) + list(G) + ([PAR] if PAR is not None else []))
        if self.sediment_storage is not None and self.sediment_storage.dim() == 3:
            _check(library().obme_add_tendencies_with_sediment(
                self.handle, _ptr_array(tracers), _ptr_array(G), _ptr(PAR), _ptr(T), _ptr(S),
                _ptr(self.sediment_storage), t, _stream()))
            return
        _check(library().obme_add_tendencies(self.handle, _ptr_array(tracers), _ptr_array(G), _ptr(PAR),
                                             _ptr(T), _ptr(S), t, _stream()))

    def step(self, tracers, nsteps, dt, t0=0.0, T=None, S=None):
        """nsteps fused RK3 steps in place (physics frozen); with a sediment,
        its storage field (self.sediment_storage) is advanced in the same pass."""
        self._check_fields(tracers)
        if self.sediment_storage is not None:
            _check(library().obme_step_with_sediment(self.handle, _ptr_array(tracers), _ptr(T), _ptr(S),
                                                     _ptr(self.sediment_storage), int(nsteps), float(dt), float(t0),
                                                     _stream()))
            return
        _check(library().obme_step(self.handle, _ptr_array(tracers), _ptr(T), _ptr(S), int(nsteps), float(dt),
                                   float(t0), _stream()))

    def step_noalloc(self, tracers, nsteps, dt, t0=0.0, T=None, S=None):
        """step() through obme_step_noalloc: the same fused RK3 steps with the
        per-call scratch in the library's static device array (no device
        allocation in the call)."""
        self._check_fields(tracers)
        _check(library().obme_step_noalloc(self.handle, _ptr_array(tracers), _ptr(T), _ptr(S), int(nsteps),
                                           float(dt), float(t0), _stream()))

    def configure_sweep(self, waves):
        """Waves per workgroup of the column sweep (1, 2, 4 or 8)."""
        _check(library().obme_sweep_configure(self.handle, int(waves)))

    @property
    def sweep_waves(self):
        return int(library().obme_sweep_waves(self.handle))

    def autotune(self, tracers, trial_steps, dt, T=None, S=None):
        """Times step() with every sweep launch shape on a scratch copy of
        `tracers` (which stay unchanged) and keeps the fastest; returns it."""
        self._check_fields(tracers)
        _check(library().obme_autotune(self.handle, _ptr_array(tracers), _ptr(T), _ptr(S), int(trial_steps),
                                       float(dt), _stream()))
        return self.sweep_waves

    def update_sediment_state(self, tracers, dt):
        """update_biogeochemical_state!(model, sediment) (Sediments/update_state.jl:6-17):
        one sediment time step of dt (the model's last stage Δt; none when
        dt is not finite) from the sinking flux of the tracers."""
        if self.sediment_storage is None:
            raise OBMEError("no sediment configured")
        self._check_fields(tracers)
        _check(library().obme_sediment_update_state(self.handle, _ptr_array(tracers), _ptr(self.sediment_storage),
                                                    float(dt), _stream()))

    # -- measurement --------------------------------------------------------------
    def set_kernel_timing(self, enable=True):
        """Bracket every kernel launch of this handle with HIP events."""
        _check(library().obme_set_kernel_timing(self.handle, 1 if enable else 0))

    def kernel_timing(self, which=KERNEL_SWEEP):
        """(summed ms, launches) of kernel `which` since the previous call."""
        ms = C.c_double()
        n = C.c_int64()
        _check(library().obme_kernel_timing(self.handle, int(which), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def co2_flux(self, DIC, Alk, T, S, u10=None, pCO2_air=None, mode=0, pH=None, flux=None):
        """pH (free scale) and CO2 flux on point sets of float64 CUDA tensors."""
        import torch
        n = DIC.numel()
        for t in (DIC, Alk, T, S) + ((u10,) if u10 is not None else ()):
            if t.dtype != torch.float64 or not t.is_cuda or t.numel() != n or not t.is_contiguous():
                raise OBMEError("co2_flux inputs must be contiguous float64 CUDA tensors of equal size")
        if pH is None:
            pH = torch.empty_like(DIC)
        if flux is None:
            flux = torch.empty_like(DIC)
        air = self.cfg.co2.air_concentration if pCO2_air is None else pCO2_air
        _check(library().obme_co2_flux(self.handle, _ptr(DIC), _ptr(Alk), _ptr(T), _ptr(S), _ptr(u10), float(air),
                                       _ptr(pH), _ptr(flux), n, int(mode), _stream()))
        return pH, flux


class HIPBatch:
    """Ensemble of independent NPZD columns on the GPU (obme_batch_*).

    bgc: an NPZD Biogeochemistry supplying the TwoBand coefficients and every
    parameter not given per column; offsets: host int array (n_columns + 1);
    params: host array [OBME_BP_COUNT, n_columns] (NaN rows = model default);
    tracers passed to step() are 4 CUDA tensors (N, P, Z, D) of offsets[-1]
    elements.  With bgc's InstantRemineralisation sediment (sediment_model),
    one sediment per column: its storage is self.sediment_storage
    (n_columns doubles), advanced by step()."""

    def __init__(self, bgc, offsets, params, dz, temperature, precision="fp32"):
        import torch
        if not torch.cuda.is_available():
            raise OBMEError("HIPBatch needs a GPU")
        self.precision = abi.OBME_FP64 if precision == "fp64" else abi.OBME_FP32_MIXED
        self.dtype = torch.float64 if self.precision == abi.OBME_FP64 else torch.float32
        self.cfg = bgc.config(self.precision)
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self.n_columns = len(off) - 1
        self.n_cells = int(off[-1])
        self._params = []
        bp = abi.BatchParams()
        params = np.asarray(params, dtype=np.float64)
        for f in range(params.shape[0]):
            if np.all(np.isnan(params[f])):
                continue
            t = torch.from_numpy(np.ascontiguousarray(params[f])).to("cuda")
            self._params.append(t)
            bp.fields[f] = t.data_ptr()
        torch.cuda.synchronize()
        h = C.c_void_p()
        _check(library().obme_batch_create(C.byref(self.cfg), self.n_columns,
                                           off.ctypes.data_as(C.POINTER(C.c_int32)), float(dz), C.byref(bp),
                                           float(temperature), C.byref(h)))
        self.handle = h
        self.sediment_storage = None
        if getattr(bgc, "sediment", None) is not None:
            self.sediment_storage = torch.zeros(self.n_columns, dtype=torch.float64, device="cuda")

    def close(self):
        if getattr(self, "handle", None):
            library().obme_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_fields(self, tracers):
        if len(tracers) != 4:
            raise OBMEError("the ensemble has 4 tracers (N, P, Z, D)")
        for t in tracers:
            if t.dtype != self.dtype or t.numel() != self.n_cells or not t.is_cuda or not t.is_contiguous():
                raise OBMEError(f"tracers must be contiguous {self.dtype} CUDA tensors of {self.n_cells} elements")

    def step(self, tracers, nsteps, dt, t0=0.0):
        """nsteps RK3 steps of every column in place (with the sediment, its
        storage advanced in the same launches)."""
        self._check_fields(tracers)
        if self.sediment_storage is not None:
            _check(library().obme_batch_step_with_sediment(self.handle, _ptr_array(tracers),
                                                           _ptr(self.sediment_storage), int(nsteps), float(dt),
                                                           float(t0), _stream()))
            return
        _check(library().obme_batch_step(self.handle, _ptr_array(tracers), int(nsteps), float(dt), float(t0),
                                         _stream()))

    def configure(self, plan, steps_per_launch=0):
        _check(library().obme_batch_configure(self.handle, int(plan), int(steps_per_launch)))

    def autotune(self, tracers, trial_steps, dt):
        self._check_fields(tracers)
        _check(library().obme_batch_autotune(self.handle, _ptr_array(tracers), int(trial_steps), float(dt),
                                             _stream()))

    def describe(self):
        buf = C.create_string_buffer(512)
        n = library().obme_batch_describe(self.handle, buf, 512)
        if n < 0:
            raise OBMEError("obme_batch_describe failed")
        return buf.value.decode()


class HIPPISCES:
    """PISCES on the GPU (obme_pisces_*) for an oceanbiome.pisces.PISCES
    description.  Tracers: 26 CUDA tensors (tracer order, parent arrays; the
    last two, T and S, are read only)."""

    def __init__(self, model, precision="fp64"):
        import torch
        if not torch.cuda.is_available():
            raise OBMEError("HIPPISCES needs a GPU")
        self.model = model
        self.precision = abi.OBME_FP64 if precision == "fp64" else abi.OBME_FP32_MIXED
        self.dtype = torch.float64 if self.precision == abi.OBME_FP64 else torch.float32
        self.cfg = model.config(self.precision)
        self._surface = None
        if self.cfg.surface_mode == abi.OBME_SURFACE_ARRAY:
            self._surface = torch.from_numpy(model.surface_PAR).to(device="cuda", dtype=self.dtype).contiguous()
            self.cfg.surface_array = self._surface.data_ptr()
        h = C.c_void_p()
        _check(library().obme_pisces_create(C.byref(self.cfg), C.byref(h)))
        self.handle = h
        # PISCES(; sediment): the InstantRemineralisation storage (Ny, Nx)
        self.sediment_storage = None
        if getattr(model, "sediment", None) is not None:
            g = model.grid
            self.sediment_storage = torch.zeros((g.Ny, g.Nx), dtype=torch.float64, device="cuda")

    def close(self):
        if getattr(self, "handle", None):
            library().obme_pisces_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_fields(self, fields, n):
        g = self.model.grid
        if len(fields) != n:
            raise OBMEError(f"expected {n} fields")
        for f in fields:
            if f.dtype != self.dtype or tuple(f.shape) != g.parent_shape or not f.is_cuda or not f.is_contiguous():
                raise OBMEError(f"field must be a contiguous {self.dtype} CUDA tensor of shape {g.parent_shape}")

    def update_biogeochemical_state(self, tracers, t=0.0):
        """(PAR, [PAR1, PAR2, PAR3], Omega, euphotic_depth, mixed_layer_PAR):
        3-D parent arrays and 2-D (Ny, Nx) tensors."""
        import torch
        self._check_fields(tracers, 26)
        g = self.model.grid
        mk3 = lambda: torch.zeros(g.parent_shape, dtype=self.dtype, device="cuda")
        PAR, bands, Om = mk3(), [mk3() for _ in range(3)], mk3()
        zeu = torch.zeros((g.Ny, g.Nx), dtype=self.dtype, device="cuda")
        ml = torch.zeros((g.Ny, g.Nx), dtype=self.dtype, device="cuda")
        _check(library().obme_pisces_update_state(self.handle, _ptr_array(tracers), _ptr(PAR), _ptr_array(bands),
                                                  _ptr(Om), _ptr(zeu), _ptr(ml), float(t), _stream()))
        return PAR, bands, Om, zeu, ml

    def update_tendencies(self, tracers, G, t=0.0):
        self._check_fields(tracers, 26)
        self._check_fields(G, 24)
        _check(library().obme_pisces_add_tendencies(self.handle, _ptr_array(tracers), _ptr_array(G), float(t),
                                                    _stream()))

    def update_sediment_state(self, tracers, dt):
        """update_biogeochemical_state!(model, sediment) (Sediments/update_state.jl:6-17):
        one sediment step of dt (none when dt is not finite) from the sinking
        flux of the tracers, on self.sediment_storage."""
        if self.sediment_storage is None:
            raise OBMEError("no sediment configured")
        self._check_fields(tracers, 26)
        _check(library().obme_pisces_sediment_update_state(self.handle, _ptr_array(tracers),
                                                           _ptr(self.sediment_storage), float(dt), _stream()))

    def step(self, tracers, nsteps, dt, t0=0.0):
        """nsteps RK3 steps in place (with the sediment, its storage advanced
        in the same passes)."""
        self._check_fields(tracers, 26)
        if self.sediment_storage is not None:
            _check(library().obme_pisces_step_with_sediment(self.handle, _ptr_array(tracers),
                                                            _ptr(self.sediment_storage), int(nsteps), float(dt),
                                                            float(t0), _stream()))
            return
        _check(library().obme_pisces_step(self.handle, _ptr_array(tracers), int(nsteps), float(dt), float(t0),
                                          _stream()))


def abi_layout_mismatches():
    """Names of the ABI structs whose ctypes size differs from the library's."""
    L = library()
    bad = []
    for idx, (name, st) in enumerate(abi.ABI_STRUCTS):
        if L.obme_abi_sizeof(idx) != C.sizeof(st):
            bad.append((name, L.obme_abi_sizeof(idx), C.sizeof(st)))
    return bad
