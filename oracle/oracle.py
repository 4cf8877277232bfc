"""ctypes wrapper of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle of the HIP library.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oceanbiome.jl_amd"))

from oceanbiome import abi  # noqa: E402  (ABI structs shared with the library)

LIB_PATH = os.path.join(HERE, "liboracle.so")

D = C.c_double
I = C.c_int
DP = C.POINTER(C.c_double)


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


class Constants(C.Structure):
    """orc_constants: equilibrium constants at (T, S[, P])."""
    _fields_ = [(n, C.c_double) for n in ("K0", "K1", "K2", "KB", "KW", "KS", "KF", "KP1", "KP2", "KP3",
                                          "KSi", "Is")]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    _lib = load_library(LIB_PATH)
    return _lib


def load_library(path):
    """An oracle library at `path` with its function signatures set (the
    committed oracle, or a variant the tests compile from its sources)."""
    L = C.CDLL(path)
    sig = {
        "orc_julia_min": (D, [D, D]),
        "orc_julia_max": (D, [D, D]),
        "orc_julia_eps": (D, [D]),
        "orc_pow_int": (D, [D, I]),
        "orc_teos10_density": (D, [D, D, D]),
        "orc_pressure_correction": (D, [I, D, D]),
        "orc_ksp_calcite": (D, [D, D, I, D]),
        "orc_ksp_aragonite": (D, [D, D, I, D]),
        "orc_virial_B": (D, [D]),
        "orc_virial_delta": (D, [D]),
        "orc_damped_newton_test": (D, [I, D, D, I, D, I, C.POINTER(I)]),
        "orc_carbon_chemistry": (D, [D, D, D, D, D, I, C.POINTER(abi.CarbonSolver), I, D, D, C.POINTER(I)]),
        "orc_carbonate_concentration": (D, [D, D, D, D, I, D, C.POINTER(abi.CarbonSolver)]),
        "orc_calcite_saturation": (D, [D, D, D, D, I, D, C.POINTER(abi.CarbonSolver)]),
        "orc_calcite_saturation_si": (D, [D, D, D, D, I, D, D, C.POINTER(abi.CarbonSolver)]),
        "orc_calcite_saturation_batch": (None, [DP, DP, DP, DP, DP, DP, DP, C.c_long, C.POINTER(abi.CarbonSolver),
                                                I]),
        "orc_schmidt_co2": (D, [C.POINTER(abi.CO2Exchange), D]),
        "orc_transfer_velocity": (D, [C.POINTER(abi.CO2Exchange), D, D, D]),
        "orc_co2_flux": (D, [C.POINTER(abi.CO2Exchange), C.POINTER(abi.CarbonSolver), D, D, D, D, D, D, I,
                             DP, C.POINTER(I)]),
        "orc_co2_flux_batch": (None, [C.POINTER(abi.CO2Exchange), C.POINTER(abi.CarbonSolver), DP, DP, DP, DP,
                                      DP, D, DP, DP, C.c_long, I]),
        "orc_npd_tendencies": (None, [I, I, C.POINTER(abi.NPDParams), DP, D, DP]),
        "orc_npd_ntracers": (I, [I, I]),
        "orc_o2_flux": (D, [C.POINTER(abi.O2Exchange), D, D, D]),
        "orc_npd_tendencies_oxygen": (None, [I, I, C.POINTER(abi.NPDParams), C.POINTER(abi.Oxygen), DP, D, DP]),
        "orc_pow_int4": (D, [D]),
        "orc_two_band_column": (None, [C.POINTER(abi.TwoBandPAR), I, DP, DP, DP, D, DP]),
        "orc_default_surface_PAR": (D, [D]),
        "orc_sinking_profile": (None, [C.POINTER(abi.Sinking), D, I, DP]),
        "orc_sinking_divergence": (None, [I, I, DP, DP, DP, DP]),
        "orc_sinking_face_value": (C.c_double, [I, I, DP, I]),
        "orc_equilibrium_constants": (None, [D, D, I, D, C.POINTER(Constants)]),
        "orc_grid_run": (I, [C.POINTER(abi.Config), C.POINTER(DP), DP, DP, DP, I, D, D, I, I]),
        "orc_grid_run_columns": (I, [C.POINTER(abi.Config), C.POINTER(DP), DP, DP, DP, I, D, D,
                                     C.POINTER(I), C.POINTER(I), I, I, I]),
        "orc_grid_update_state": (I, [C.POINTER(abi.Config), C.POINTER(DP), DP, DP, D]),
        "orc_scale_negatives": (None, [C.POINTER(abi.NegativeScaling), DP]),
        "orc_grid_scale_negatives": (I, [C.POINTER(abi.Config), C.POINTER(DP)]),
        "orc_grid_tendencies": (I, [C.POINTER(abi.Config), C.POINTER(DP), C.POINTER(DP), DP, DP, DP, D, DP]),
        "orc_grid_run_sediment": (I, [C.POINTER(abi.Config), C.POINTER(DP), DP, DP, DP, I, D, D, I, I, DP, DP, DP,
                                      DP, C.POINTER(C.c_long)]),
        "orc_grid_sediment_update_state": (I, [C.POINTER(abi.Config), C.POINTER(DP), DP, DP, DP, DP,
                                               C.POINTER(C.c_long), D]),
        "orc_sediment_flux": (D, [C.POINTER(abi.Config), DP, I]),
        "orc_burial_efficiency": (D, [C.POINTER(abi.Sediment), D]),
        "orc_sediment_fields": (I, [C.POINTER(abi.Sediment)]),
        "orc_sediment_tendencies": (None, [C.POINTER(abi.Sediment), DP, D, DP]),
        "orc_smg_coupled": (None, [C.POINTER(abi.Sediment), DP, D, D, D, DP, DP, DP]),
        "orc_grid_tendencies_sediment": (I, [C.POINTER(abi.Config), C.POINTER(DP), C.POINTER(DP), DP, DP, DP, D, DP,
                                             DP]),
        "orc_batch_run": (I, [C.POINTER(abi.Config), I, C.POINTER(I), C.POINTER(DP), D, D, C.POINTER(DP), I, D,
                              D, I]),
        "orc_batch_run_sediment": (I, [C.POINTER(abi.Config), I, C.POINTER(I), C.POINTER(DP), D, D, C.POINTER(DP),
                                       I, D, D, I, DP, DP, DP, C.POINTER(D), C.POINTER(C.c_long)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def _ptrs(arrs):
    return (DP * len(arrs))(*[_dp(a) for a in arrs])


# ---- scalar helpers -------------------------------------------------------

def julia_min(x, y):
    return lib().orc_julia_min(x, y)


def julia_max(x, y):
    return lib().orc_julia_max(x, y)


def julia_eps(x):
    return lib().orc_julia_eps(x)


def pow_int(x, n):
    return lib().orc_pow_int(x, n)


def density(T, S, Pbar):
    return lib().orc_teos10_density(T, S, Pbar)


def equilibrium_constants(Tk, S, P=None):
    out = Constants()
    lib().orc_equilibrium_constants(Tk, S, 0 if P is None else 1, 0.0 if P is None else P, C.byref(out))
    return out


def ksp_calcite(Tk, S, P=None):
    return lib().orc_ksp_calcite(Tk, S, 0 if P is None else 1, 0.0 if P is None else P)


def ksp_aragonite(Tk, S, P=None):
    return lib().orc_ksp_aragonite(Tk, S, 0 if P is None else 1, 0.0 if P is None else P)


def virial_B(Tk):
    return lib().orc_virial_B(Tk)


def virial_delta(Tk):
    return lib().orc_virial_delta(Tk)


def schmidt_co2(gas, T):
    return lib().orc_schmidt_co2(C.byref(gas), T)


def transfer_velocity(gas, u10, T, S):
    return lib().orc_transfer_velocity(C.byref(gas), u10, T, S)


def default_surface_PAR(t):
    return lib().orc_default_surface_PAR(t)


def sinking_profile(speed, Nz, open_bottom=True, smoothing_distance=2.0):
    s = abi.Sinking()
    s.open_bottom = 1 if open_bottom else 0
    s.smoothing_distance = smoothing_distance
    w = np.zeros(Nz + 1)
    lib().orc_sinking_profile(C.byref(s), speed, Nz, _dp(w))
    return w


def sinking_divergence(w, c, dz, scheme=None):
    Nz = len(c)
    out = np.zeros(Nz)
    lib().orc_sinking_divergence(abi.OBME_UPWIND1 if scheme is None else scheme, Nz,
                                 _dp(np.ascontiguousarray(w, dtype=np.float64)),
                                 _dp(np.ascontiguousarray(c, dtype=np.float64)),
                                 _dp(np.ascontiguousarray(dz, dtype=np.float64)), _dp(out))
    return out


def sinking_face_value(c, k, scheme):
    c = np.ascontiguousarray(c, dtype=np.float64)
    return lib().orc_sinking_face_value(scheme, len(c), _dp(c), int(k))


def pressure_correction(which, Tk, P):
    return lib().orc_pressure_correction(which, Tk, P)


def damped_newton_test(which, x0, damping=0.5, lower=None, fast_exit=0):
    it = I(0)
    x = lib().orc_damped_newton_test(which, x0, damping, 0 if lower is None else 1,
                                     0.0 if lower is None else lower, fast_exit, C.byref(it))
    return x, it.value


def carbon_chemistry(solver, DIC, Alk, T, S, pH=None, output="fCO2", fast_exit=1, silicate=0.0,
                     phosphate=0.0):
    outs = {"fCO2": 0, "pCO2": 1, "pHf": 2, "pHt": 3, "pHs": 4}
    ev = I(0)
    v = lib().orc_carbon_chemistry(DIC, Alk, T, S, -1000.0 if pH is None else pH, outs[output],
                                   C.byref(solver), fast_exit, silicate, phosphate, C.byref(ev))
    return v, ev.value


def co2_flux(gas, solver, DIC, Alk, T, S, u10, pCO2_air, fast_exit=1):
    pH = D(0)
    ev = I(0)
    F = lib().orc_co2_flux(C.byref(gas), C.byref(solver), DIC, Alk, T, S, u10, pCO2_air, fast_exit,
                           C.byref(pH), C.byref(ev))
    return F, pH.value, ev.value


def o2_schmidt(o2, T):
    """OxygenPolynomialSchmidtNumber (GasExchange/schmidt_number.jl:15-16) as
    the quartic PolynomialParameterisation{4} (generic_parameterisations.jl:31-32)."""
    c = o2.schmidt
    return c[0] + c[1] * T + c[2] * (T * T) + c[3] * (T * T * T) + c[4] * lib().orc_pow_int4(T)


def o2_air_concentration(o2, T, S):
    """PartiallySolubleGas surface value (gas_solubility.jl:23-24): the air
    concentration times the Wanninkhof (1992) solubility β/Tk (:37-48; the
    salinity polynomial uses B2 twice, as written there)."""
    a = o2.solubility
    Tk = T + 273.15
    Tk100 = Tk / 100
    beta = math.exp(a[0] + a[1] / Tk100 + a[2] * math.log(Tk100) + S * (a[3] + a[4] * Tk100 + a[4] * Tk100 ** 2))
    return o2.air_concentration * (beta / Tk)


def o2_flux(o2, O2, T, S):
    """GasExchange (gas_exchange.jl:26-38) for O2: k (O2 - C_air α), k =
    k660(u) / sqrt(Sc(T) / 660) (gas_transfer_velocity.jl:33-34)."""
    v = o2.transfer_velocity
    u = o2.wind_speed
    k660 = v[0] + v[1] * u + v[2] * (u * u)
    k = k660 / math.sqrt(o2_schmidt(o2, T) / 660.0) * 1.0
    return k * (O2 - o2_air_concentration(o2, T, S))


def calcite_saturation(solver, DIC, Alk, T, S, P=None, silicate=0.0):
    return lib().orc_calcite_saturation_si(DIC, Alk, T, S, 0 if P is None else 1, 0.0 if P is None else P,
                                           silicate, C.byref(solver))


def calcite_saturation_batch(solver, DIC, Alk, T, S, P, Si, nthreads=8):
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (DIC, Alk, T, S, P, Si)]
    out = np.empty(arrs[0].size)
    lib().orc_calcite_saturation_batch(*[_dp(a) for a in arrs], _dp(out), out.size, C.byref(solver), nthreads)
    return out.reshape(np.shape(DIC))


def co2_flux_batch(gas, solver, DIC, Alk, T, S, u10, pCO2_air, nthreads=8):
    n = DIC.size
    pH = np.empty(n)
    F = np.empty(n)
    lib().orc_co2_flux_batch(C.byref(gas), C.byref(solver), _dp(DIC), _dp(Alk), _dp(T), _dp(S), _dp(u10),
                             pCO2_air, _dp(pH), _dp(F), n, nthreads)
    return pH, F


def npd_tendencies(model, carbonate, params, vals, PAR):
    n = lib().orc_npd_ntracers(model, carbonate)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    out = np.zeros(n)
    lib().orc_npd_tendencies(model, carbonate, C.byref(params), _dp(v), PAR, _dp(out))
    return out


def npd_tendencies_oxygen(carbonate, params, oxygen, vals, PAR, model=abi.OBME_LOBSTER):
    """NPZD or LOBSTER transitions with the Oxygen component (O2 last)."""
    n = lib().orc_npd_ntracers(model, carbonate) + 1
    v = np.ascontiguousarray(vals, dtype=np.float64)
    out = np.zeros(n)
    lib().orc_npd_tendencies_oxygen(model, carbonate, C.byref(params), C.byref(oxygen), _dp(v), PAR, _dp(out))
    return out


def two_band_column(par, zc, zf, P, PAR0):
    Nz = len(zc)
    out = np.zeros(Nz)
    lib().orc_two_band_column(C.byref(par), Nz, _dp(zc), _dp(zf), _dp(np.ascontiguousarray(P)), PAR0,
                              _dp(out))
    return out


# ---- grid drivers ------------------------------------------------------------

def grid_run(cfg, tracers, nsteps, dt, t0=0.0, T3=None, S3=None, PAR0=None, nthreads=8, fast_exit=1):
    """Advance Oceananigans-layout float64 arrays in place."""
    lib().orc_grid_run(C.byref(cfg), _ptrs(tracers), _dp(T3), _dp(S3), _dp(PAR0), nsteps, dt, t0,
                       nthreads, fast_exit)


class SedimentState:
    """A BiogeochemicalSediment's state on the oracle side: its fields and
    their time stepper's Gⁿ, G⁻ ((Ny, Nx) for one field, as
    InstantRemineralisation's storage; (fields, Ny, Nx) for more, as
    SimpleMultiG's Ns, Nf, Nr), and its clock (last Δt, iteration)."""

    def __init__(self, Nx, Ny, fields=1):
        shape = (Ny, Nx) if fields == 1 else (fields, Ny, Nx)
        self.storage = np.zeros(shape)
        self.Gn = np.zeros(shape)
        self.Gm = np.zeros(shape)
        self.last_dt = math.inf
        self.iteration = 0


def grid_run_sediment(cfg, tracers, state, nsteps, dt, t0=0.0, T3=None, S3=None, PAR0=None, nthreads=8,
                      fast_exit=1):
    """grid_run with the sediment: after every substep, the sediment's time
    step of the stage Δt (update_biogeochemical_state!, Sediments/update_state.jl:6-17)."""
    last = C.c_double(state.last_dt)
    it = C.c_long(state.iteration)
    lib().orc_grid_run_sediment(C.byref(cfg), _ptrs(tracers), _dp(T3), _dp(S3), _dp(PAR0), nsteps, dt, t0,
                                nthreads, fast_exit, _dp(state.storage), _dp(state.Gn), _dp(state.Gm),
                                C.byref(last), C.byref(it))
    state.last_dt = last.value
    state.iteration = it.value


def grid_sediment_update_state(cfg, tracers, state, dt):
    """update_biogeochemical_state!(model, sediment): one sediment step of dt."""
    last = C.c_double(state.last_dt)
    it = C.c_long(state.iteration)
    lib().orc_grid_sediment_update_state(C.byref(cfg), _ptrs(tracers), _dp(state.storage), _dp(state.Gn),
                                         _dp(state.Gm), C.byref(last), C.byref(it), dt)
    state.last_dt = last.value
    state.iteration = it.value


def sediment_flux(cfg, U):
    """The sinking flux through the bottom face of one column, U[n, k]."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    return lib().orc_sediment_flux(C.byref(cfg), _dp(U), U.shape[1])


def burial_efficiency(sediment, f):
    return lib().orc_burial_efficiency(C.byref(sediment), float(f))


def sediment_fields(sediment):
    return lib().orc_sediment_fields(C.byref(sediment))


def sediment_tendencies(sediment, u, f):
    """The tendencies of the sediment's fields u from the tracked flux f."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    G = np.zeros(3)
    lib().orc_sediment_tendencies(C.byref(sediment), _dp(u), float(f), _dp(G))
    return G[:sediment_fields(sediment)]


def smg_coupled(sediment, u, NO3, NH4, O2):
    """SimpleMultiG's coupled tendencies (NH4, NO3, O2), per unit area."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = [C.c_double(), C.c_double(), C.c_double()]
    lib().orc_smg_coupled(C.byref(sediment), _dp(u), float(NO3), float(NH4), float(O2), *[C.byref(o) for o in out])
    return tuple(o.value for o in out)


def grid_run_columns(cfg, tracers, cols_i, cols_j, nsteps, dt, t0=0.0, T3=None, S3=None, PAR0=None,
                     nthreads=8, fast_exit=1):
    ci = np.ascontiguousarray(cols_i, dtype=np.int32)
    cj = np.ascontiguousarray(cols_j, dtype=np.int32)
    lib().orc_grid_run_columns(C.byref(cfg), _ptrs(tracers), _dp(T3), _dp(S3), _dp(PAR0), nsteps, dt, t0,
                               ci.ctypes.data_as(C.POINTER(I)), cj.ctypes.data_as(C.POINTER(I)), len(ci),
                               nthreads, fast_exit)


def scale_negatives(scaling, values):
    """ScaleNegativeTracers on one cell: `values` indexed by tracer; returns the new values."""
    v = np.ascontiguousarray(values, dtype=np.float64).copy()
    lib().orc_scale_negatives(C.byref(scaling), _dp(v))
    return v


def grid_scale_negatives(cfg, tracers):
    """The scaling modifier in place on Oceananigans-layout float64 arrays."""
    lib().orc_grid_scale_negatives(C.byref(cfg), _ptrs(tracers))


def grid_update_state(cfg, tracers, PAR, t=0.0, PAR0=None):
    lib().orc_grid_update_state(C.byref(cfg), _ptrs(tracers), _dp(PAR0), _dp(PAR), t)


def grid_tendencies(cfg, tracers, G, t=0.0, T3=None, S3=None, PAR0=None, PAR=None, sediment_fields=None):
    """Tendencies added into G; sediment_fields: the sediment's fields (a
    SimpleMultiG's (3, Ny, Nx)) for its coupled tendencies, or None."""
    fields = None if sediment_fields is None else np.ascontiguousarray(sediment_fields, dtype=np.float64)
    lib().orc_grid_tendencies_sediment(C.byref(cfg), _ptrs(tracers), _ptrs(G), _dp(T3), _dp(S3), _dp(PAR0), t,
                                       _dp(PAR), _dp(fields))


def batch_run(cfg, offsets, params, dz, T_const, U, nsteps, dt, t0=0.0, nthreads=8, sediment=None):
    """The ensemble's columns, each run as one column; with `sediment` (a
    SedimentState of one field over the columns: storage of shape (1,
    n_columns)) the InstantRemineralisation sediment of cfg under each."""
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    if sediment is None:
        lib().orc_batch_run(C.byref(cfg), len(off) - 1, off.ctypes.data_as(C.POINTER(I)), _ptrs(params), dz,
                            T_const, _ptrs(U), nsteps, dt, t0, nthreads)
        return
    last = C.c_double(sediment.last_dt)
    it = C.c_long(sediment.iteration)
    lib().orc_batch_run_sediment(C.byref(cfg), len(off) - 1, off.ctypes.data_as(C.POINTER(I)), _ptrs(params), dz,
                                 T_const, _ptrs(U), nsteps, dt, t0, nthreads, _dp(sediment.storage),
                                 _dp(sediment.Gn), _dp(sediment.Gm), C.byref(last), C.byref(it))
    sediment.last_dt = last.value
    sediment.iteration = it.value
