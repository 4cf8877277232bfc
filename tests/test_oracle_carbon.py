"""Pins of the FP64 oracle's carbonate chemistry, CO2 exchange and damped
Newton solver against the known answers the reference holds
(tests/golden/reference_values.json) and against what the mathematics fixes.
CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from oceanbiome.models import CarbonChemistry, CarbonDioxideGasExchangeBoundaryCondition, DampedNewtonRaphsonSolver

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))
JULIA_RTOL = math.sqrt(np.finfo(float).eps)   # Julia's default isapprox rtol


def solver():
    cc = CarbonChemistry()
    return cc.solver.struct(cc.initial_pH_guess)


# ---- full chain: density, constants, Newton solve, virial pCO2 -------------

@pytest.mark.parametrize("case", REF["doctests"]["cases"])
def test_doctests_full_precision(case):
    """carbon_chemistry.jl:55-62 doctests: agreement to a few ulp (same libm
    calls; the printed Julia values carry 16-17 significant digits)."""
    v, _ = O.carbon_chemistry(solver(), case["DIC"], case.get("Alk", 0.0), case["T"], case["S"],
                              pH=case.get("pH"), output=case["output"])
    assert abs(v - case["expected"]) <= 4 * np.finfo(float).eps * abs(case["expected"]), (v, case)


def test_equilibrium_constants_dickson():
    e = REF["equilibrium_constants"]
    k = O.equilibrium_constants(e["Tk"], e["S"])
    for name, (val, atol) in e["log"].items():
        assert abs(math.log(getattr(k, name)) - val) <= atol, name
    for name, (val, atol) in e["log10"].items():
        if name.startswith("KSP"):
            x = O.ksp_calcite(e["Tk"], e["S"]) if name.endswith("calcite") else O.ksp_aragonite(e["Tk"], e["S"])
        else:
            x = getattr(k, name)
        assert abs(math.log10(x) - val) <= atol, name


def test_pressure_corrections_zeebe():
    e = REF["equilibrium_constants"]
    which = {"K1": 1, "K2": 2, "KB": 3, "KW": 4, "KS": 5, "KF": 6, "KP1": 7, "KP2": 8, "KP3": 9,
             "KSP_calcite": 10, "KSP_aragonite": 11}
    for name, (val, atol) in e["pressure_correction"].items():
        assert abs(O.pressure_correction(which[name], e["Tk"], e["P"]) - val) <= atol, name


def test_pressure_applied_to_constants():
    """With pressure, every constant is its surface value times its correction
    factor (equilibrium_constants.jl:11-40 applied at :118, :164, ...)."""
    Tk, S, P = 283.15, 34.0, 150.0
    k0 = O.equilibrium_constants(Tk, S)
    kp = O.equilibrium_constants(Tk, S, P)
    for name, w in (("K1", 1), ("K2", 2), ("KB", 3), ("KW", 4), ("KP1", 7), ("KP2", 8), ("KP3", 9)):
        ratio = getattr(kp, name) / getattr(k0, name)
        assert ratio == pytest.approx(O.pressure_correction(w, Tk, P), rel=1e-14), name
    assert O.pressure_correction(1, Tk, 0.0) == 1.0


def test_gas_exchange_constants():
    g = REF["gas_exchange"]
    gas = CarbonDioxideGasExchangeBoundaryCondition().struct()
    val, atol = g["schmidt_CO2_20C"]
    assert abs(O.schmidt_co2(gas, 20.0) - val) <= atol
    val, atol = g["virial_B_298_15K"]
    assert abs(O.virial_B(298.15) - val) <= atol
    val, atol = g["virial_delta_298_15K"]
    assert abs(O.virial_delta(298.15) - val) <= atol


def test_pCO2_350():
    c = REF["gas_exchange"]["pCO2_case"]
    v, _ = O.carbon_chemistry(solver(), c["DIC"], c["Alk"], c["T"], c["S"], output="pCO2")
    assert abs(v - c["expected"]) <= c["atol"]


def test_co2_flux_magnitude_and_sign():
    c = REF["gas_exchange"]["flux_case"]
    gas = CarbonDioxideGasExchangeBoundaryCondition().struct()
    F, pH, _ = O.co2_flux(gas, solver(), c["DIC"], c["Alk"], c["T"], c["S"], gas.wind_speed,
                          c["air_concentration"])
    assert abs(F - c["expected"]) <= c["atol"]
    assert 7.0 < pH < 9.0


def test_flux_is_k_times_pCO2_difference():
    """F = k (pCO2_w - pCO2_a) (gas_exchange.jl:26-38) with k from the Schmidt
    scaled transfer velocity (gas_transfer_velocity.jl:24-34): the flux is
    affine in pCO2_air with slope -k, and zero at pCO2_air = pCO2_w."""
    gas = CarbonDioxideGasExchangeBoundaryCondition().struct()
    s = solver()
    DIC, Alk, T, S, u = 2100.0, 2350.0, 12.0, 34.5, 7.0
    F1, _, _ = O.co2_flux(gas, s, DIC, Alk, T, S, u, 400.0)
    F2, _, _ = O.co2_flux(gas, s, DIC, Alk, T, S, u, 500.0)
    k = O.transfer_velocity(gas, u, T, S)
    assert (F1 - F2) / 100.0 == pytest.approx(k, rel=1e-10)
    pw, _ = O.carbon_chemistry(s, DIC, Alk, T, S, output="pCO2")
    F0, _, _ = O.co2_flux(gas, s, DIC, Alk, T, S, u, pw)
    assert abs(F0) <= 1e-12 * abs(F1)


def test_transfer_velocity_ho06():
    """Ho06 (gas_transfer_velocity.jl:61): k660 = 0.266 u^2 cm/h, scaled by
    sqrt(660 / Sc) (:24-34), times the solubility K0 rho(T, S, 0) / 10^3
    (gas_solubility.jl:56-65) that converts the pCO2 difference to mmol/m^3."""
    gas = CarbonDioxideGasExchangeBoundaryCondition().struct()
    for u, T, S in ((2.0, 15.0, 35.0), (10.0, 3.0, 33.0), (0.5, 28.0, 37.0)):
        Sc = O.schmidt_co2(gas, T)
        k660 = 0.266 * u ** 2 / 3600 / 100
        sol = O.equilibrium_constants(T + 273.15, S).K0 * O.density(T, S, 0.0) / 1000
        assert O.transfer_velocity(gas, u, T, S) == pytest.approx(k660 * math.sqrt(660 / Sc) * sol, rel=1e-12)
    # quadratic in the wind speed
    assert O.transfer_velocity(gas, 10.0, 15.0, 35.0) == pytest.approx(25 * O.transfer_velocity(gas, 2.0, 15.0, 35.0),
                                                                      rel=1e-13)


def test_pH_given_consistency():
    """Solving for pH and then evaluating with that pH gives the same pCO2
    (carbon_chemistry.jl:155-162: the pH branch skips the solve)."""
    s = solver()
    for DIC, Alk, T, S in ((2000.0, 2200.0, 5.0, 33.0), (2150.0, 2400.0, 25.0, 36.0)):
        pH, _ = O.carbon_chemistry(s, DIC, Alk, T, S, output="pHf")
        p1, _ = O.carbon_chemistry(s, DIC, Alk, T, S, output="fCO2")
        p2, _ = O.carbon_chemistry(s, DIC, Alk, T, S, pH=pH, output="fCO2")
        assert p2 == pytest.approx(p1, rel=1e-9)


def test_pH_scales_ordered():
    """Total and seawater scales add sulfate (and fluoride) to the free scale:
    pH_free > pH_total > pH_seawater (carbon_chemistry.jl:195-211)."""
    s = solver()
    f, _ = O.carbon_chemistry(s, 2100.0, 2300.0, 15.0, 35.0, output="pHf")
    t, _ = O.carbon_chemistry(s, 2100.0, 2300.0, 15.0, 35.0, output="pHt")
    w, _ = O.carbon_chemistry(s, 2100.0, 2300.0, 15.0, 35.0, output="pHs")
    assert f > t > w
    assert 0.05 < f - t < 0.2


def test_fast_exit_bit_identical():
    """Stopping the damped Newton loop at an exact fixed point gives the
    result of the full loop bit for bit (once x stops changing, every later
    iteration repeats the same arithmetic)."""
    rng = np.random.default_rng(7)
    s = solver()
    for _ in range(200):
        DIC = rng.uniform(1900, 2300)
        Alk = rng.uniform(2100, 2500)
        T = rng.uniform(-2, 32)
        S = rng.uniform(30, 38)
        a, ea = O.carbon_chemistry(s, DIC, Alk, T, S, output="pHf", fast_exit=0)
        b, eb = O.carbon_chemistry(s, DIC, Alk, T, S, output="pHf", fast_exit=1)
        assert a == b
        assert eb <= ea


def test_flux_batch_matches_scalar():
    rng = np.random.default_rng(3)
    n = 64
    DIC = rng.uniform(1900, 2300, n)
    Alk = rng.uniform(2100, 2500, n)
    T = rng.uniform(-2, 32, n)
    S = rng.uniform(30, 38, n)
    u = 10 * np.exp(0.3 * rng.standard_normal(n))
    gas = CarbonDioxideGasExchangeBoundaryCondition().struct()
    s = solver()
    pH, F = O.co2_flux_batch(gas, s, DIC, Alk, T, S, u, 420.0, nthreads=2)
    for q in range(n):
        f1, p1, _ = O.co2_flux(gas, s, DIC[q], Alk[q], T[q], S[q], u[q], 420.0)
        assert F[q] == f1 and pH[q] == p1


# ---- damped Newton on the reference's analytic test problems ---------------

def _dn(which, x0):
    x, _ = O.damped_newton_test(which, float(x0), damping=REF["damped_newton"]["damping"])
    return x


def test_damped_newton_exact_roots():
    for fn, spec in REF["damped_newton"]["exact"].items():
        which = int(fn[1:])
        got = [_dn(which, x0) == spec["root"] for x0 in spec["x0"]]
        assert got == spec["expect"], (fn, got)


def test_damped_newton_approx_roots():
    for fn, spec in REF["damped_newton"]["approx"].items():
        which = int(fn[1:])
        got = [abs(_dn(which, x0) - r) <= JULIA_RTOL * abs(r) for x0, r in zip(spec["x0"], spec["root"])]
        assert got == spec["expect"], (fn, got)


def test_damped_newton_respects_lower_bound():
    """bounded_λ (solvers.jl:123-141) keeps iterates above the lower bound:
    f2 from x0 = 0.5 with lower bound 0 converges to +sqrt(2)."""
    x, _ = O.damped_newton_test(2, 0.5, damping=0.5, lower=0.0)
    assert x == pytest.approx(math.sqrt(2), rel=1e-14)


# ---- Julia scalar semantics -------------------------------------------------

def test_julia_min_max_nan_and_signed_zero():
    assert math.isnan(O.julia_min(float("nan"), 1.0))
    assert math.isnan(O.julia_min(1.0, float("nan")))
    assert math.isnan(O.julia_max(float("nan"), 1.0))
    assert math.copysign(1.0, O.julia_max(-0.0, 0.0)) == 1.0
    assert math.copysign(1.0, O.julia_min(-0.0, 0.0)) == -1.0
    assert O.julia_min(2.0, 3.0) == 2.0 and O.julia_max(2.0, 3.0) == 3.0


def test_julia_eps_and_integer_powers():
    assert O.julia_eps(1.0) == np.finfo(float).eps
    assert O.julia_eps(0.0) == 5e-324
    assert O.julia_eps(2.0) == 2 * np.finfo(float).eps
    assert O.pow_int(3.0, 2) == 9.0
    assert O.pow_int(2.0, 10) == 1024.0
    assert O.pow_int(0.1, 10) == pytest.approx(1e-10, rel=1e-15)


def test_density_polynomial_height_convention():
    """teos10_polynomial_approximation (seawater_density.jl:34-40) passes
    Z = +10 Pbar as the geopotential height of the polynomial: sea water at
    10 C, 35 psu is about 1027 kg/m^3 at Pbar = 0 and, with that sign, a
    positive Pbar reads as height above the surface and lowers the density,
    by about 0.0045 kg/m^3 per metre."""
    r0 = O.density(10.0, 35.0, 0.0)
    r1 = O.density(10.0, 35.0, 1.0)
    rm = O.density(10.0, 35.0, -1.0)
    assert 1026.0 < r0 < 1028.5
    assert r1 < r0 < rm
    assert (rm - r1) / 20.0 == pytest.approx(0.0045, rel=0.1)
