"""Pins of the oracle's sinking reconstructions (oracle_npd.c face_value,
sinking_divergence): fifth-order WENO (WENO-Z weights on Jiang-Shu
smoothness indicators), its third-order reduction near the boundaries, the
centred and the upwind face values, checked against what the mathematics
fixes rather than against the formula itself:

 - every convex combination of the three-cell reconstructions is exact for
   the cell averages of a quadratic, so the WENO5 faces return the
   quadratic's value at the face whatever the weights; the WENO3 faces are
   exact for linear data only, the upwind faces return the upwind cell;
 - data mirrored about the upwind cell make the outer smoothness
   indicators equal (tau = 0), so the weights are the linear ones and the
   face value is the five-point linear scheme (2, -13, 47, 27, -3)/60 (the
   third-order faces: (-1, 5, 2)/6);
 - a step scaled so that the smoothness indicators are of the size of eps
   gives weights that depend on eps, d_r and the indicators' coefficients:
   evaluated by hand for that case;
 - the divergence telescopes: sum dz div = -w(bottom) c(bottom cell).

Each pin is shown to catch a planted mistake in a recompiled oracle.
Parity with Oceananigans' own WENO is unpinned (Oceananigans is not part of
the reference tree)."""
import ctypes as C
import shutil

import numpy as np
import pytest

import mutants
import oracle as O
from oceanbiome import abi

UP, CEN, WENO = abi.OBME_UPWIND1, abi.OBME_CENTERED2, abi.OBME_WENO5


def _face(L, c, k, scheme):
    c = np.ascontiguousarray(c, dtype=np.float64)
    return L.orc_sinking_face_value(scheme, len(c), c.ctypes.data_as(C.POINTER(C.c_double)), int(k))


def _div(L, w, c, dz, scheme):
    out = np.zeros(len(c))
    dp = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
    w, c, dz = (np.ascontiguousarray(a, dtype=np.float64) for a in (w, c, dz))
    L.orc_sinking_divergence(scheme, len(c), dp(w), dp(c), dp(dz), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def _quadratic(Nz, a=1.3, b=-0.4, q=0.07):
    """Cell averages over unit cells centred at k of a + b z + q z^2, and the
    polynomial's value at face k (z = k - 1/2)."""
    k = np.arange(Nz, dtype=np.float64)
    avg = a + b * k + q * (k * k + 1.0 / 12)
    face = lambda f: a + b * (f - 0.5) + q * (f - 0.5) ** 2
    return avg, face


def pin_weno5_quadratic(L):
    Nz = 12
    c, face = _quadratic(Nz)
    got = [_face(L, c, k, WENO) for k in range(3, Nz - 2)]
    return np.array(got), np.array([face(k) for k in range(3, Nz - 2)])


def pin_weno3_linear(L):
    Nz = 9
    c, face = _quadratic(Nz, q=0.0)
    ks = (2, Nz - 2)
    return np.array([_face(L, c, k, WENO) for k in ks]), np.array([face(k) for k in ks])


def pin_weno3_linear_weights(L):
    """Face 2 (third order) with c(1) = c(3) = A about the upwind cell 2:
    equal indicators, so the linear weights (1/3, 2/3) give (A + 5 C)/6;
    on quadratic data the nonlinear weights move it off the face value but
    keep it between the two candidates."""
    A, Cc = 0.6, 1.9
    c = np.array([0.2, A, Cc, A, 0.7, 1.3, 0.5, 0.9, 1.4])
    got = [_face(L, c, 2, WENO)]
    exp = [(A + 5 * Cc) / 6]
    q, face = _quadratic(9)
    v = _face(L, q, 2, WENO)
    q0, q1 = (-q[3] + 3 * q[2]) / 2, (q[2] + q[1]) / 2
    inside = min(q0, q1) <= v <= max(q0, q1) and abs(v - face(2)) > 1e-6
    return np.array(got + [1.0 if inside else 0.0]), np.array(exp + [1.0])


def pin_upwind_faces(L):
    Nz = 9
    c, _ = _quadratic(Nz)
    ks = (1, Nz - 1)
    got = [_face(L, c, k, WENO) for k in ks] + [_face(L, c, k, UP) for k in range(1, Nz)]
    exp = [c[k] for k in ks] + [c[k] for k in range(1, Nz)]
    return np.array(got), np.array(exp)


def pin_centred(L):
    Nz = 7
    c, _ = _quadratic(Nz)
    return np.array([_face(L, c, k, CEN) for k in range(1, Nz)]), (c[:-1] + c[1:]) / 2


def pin_linear_weights(L):
    """Mirrored data about the upwind cell k: c(k±1) = A, c(k±2) = B."""
    A, B, Cc = 0.8, 2.9, 1.7
    c = np.array([0.3, 0.9, B, A, Cc, A, B, 0.4, 1.1])
    k = 4
    return np.array([_face(L, c, k, WENO)]), np.array([(-B + 14 * A + 47 * Cc) / 60])


def pin_step_weights(L):
    """cells k-2, k-1 = 0 and k, k+1, k+2 = s (a step between k-1 and k):
    q = (s, 2s/3, s/3), beta = (0, 4/3 s^2, 10/3 s^2), tau = 10/3 s^2; with
    s = 1e-4 the indicators are of the size of eps = 1e-8."""
    s = 1e-4
    eps = 1e-8
    c = np.array([0.0, 0.0, 0.0, 0.0, s, s, s, s])
    k = 4
    b = np.array([0.0, 4.0 / 3 * s * s, 10.0 / 3 * s * s])
    tau = 10.0 / 3 * s * s
    d = np.array([0.1, 0.6, 0.3])
    alpha = d * (1 + (tau / (b + eps)) ** 2)
    q = np.array([s, 2 * s / 3, s / 3])
    return np.array([_face(L, c, k, WENO)]), np.array([np.dot(alpha, q) / alpha.sum()])


def pin_telescoping(L):
    rng = np.random.default_rng(3)
    out_got, out_exp = [], []
    for scheme in (UP, CEN, WENO):
        Nz = 17
        c = rng.uniform(0.1, 2.0, Nz)
        dz = rng.uniform(1.0, 4.0, Nz)
        w = -rng.uniform(0.5, 2.0, Nz + 1) * 1e-4
        w[Nz] = 0.0
        div = _div(L, w, c, dz, scheme)
        out_got.append(np.sum(div * dz))
        out_exp.append(-w[0] * c[0])
    return np.array(out_got), np.array(out_exp)


def pin_divergence_uses_the_face_values(L):
    """div(k) = (w(k+1) f(k+1) - w(k) f(k)) / dz(k) with f the face values."""
    rng = np.random.default_rng(4)
    Nz = 11
    c = rng.uniform(0.1, 2.0, Nz)
    dz = np.full(Nz, 2.0)
    w = np.full(Nz + 1, -3e-4)
    w[Nz] = 0.0
    f = [c[0]] + [_face(L, c, k, WENO) for k in range(1, Nz)] + [c[-1]]
    exp = [(w[k + 1] * f[k + 1] - w[k] * f[k]) / dz[k] for k in range(Nz)]
    return _div(L, w, c, dz, WENO), np.array(exp)


PINS = {n: f for n, f in globals().items() if n.startswith("pin_")}


def failing_pins(L, rtol=1e-12):
    bad = {}
    for name, f in PINS.items():
        got, exp = f(L)
        if not np.allclose(got, exp, rtol=rtol, atol=1e-15 * np.max(np.abs(exp))):
            bad[name] = (got, exp)
    return bad


def test_sinking_pins_hold_for_the_oracle():
    assert failing_pins(O.lib()) == {}


MUTATIONS = {
    "weno5_q0": ("(2 * cp2 - 7 * cp1 + 11 * c0) / 6", "(2 * cp2 - 7 * cp1 + 11 * c0) / 5"),
    "weno5_q2_sign": ("(2 * c0 + 5 * cm1 - cm2) / 6", "(2 * c0 + 5 * cm1 + cm2) / 6"),
    "weno5_linear_weights": ("0.1 * (1 + sq(tau / (b0", "0.3 * (1 + sq(tau / (b0"),
    "weno5_beta_coefficient": ("13.0 / 12 * sq(cp1 - 2 * c0 + cm1)", "13.0 / 13 * sq(cp1 - 2 * c0 + cm1)"),
    "weno5_tau": ("const double tau = fabs(b0 - b2);", "const double tau = fabs(b0 - b1);"),
    "weno_eps": ("#define WENO_EPS 1e-8", "#define WENO_EPS 1e-6"),
    "weno_z_power": ("0.6 * (1 + sq(tau / (b1 + WENO_EPS)))", "0.6 * (1 + tau / (b1 + WENO_EPS))"),
    "weno3_q0": ("(-cp1 + 3 * c0) / 2", "(-cp1 + 3 * c0) / 3"),
    # (the widened ranges stay inside the column: the mutants read no cell past its ends)
    "weno3_range": ("if (k >= 2 && k <= Nz - 2) return weno3", "if (k >= 1 && k <= Nz - 2) return weno3"),
    "weno5_range": ("if (k >= 3 && k <= Nz - 3) return weno5", "if (k >= 2 && k <= Nz - 3) return weno5"),
    "centred_face": ("return (c[k - 1] + c[k]) / 2;", "return (c[k] + c[k + 1]) / 2;"),
    "divergence_bottom_face": ("const double cf_lo = (k > 0) ? face_value(scheme, Nz, c, k) : c[k];",
                               "const double cf_lo = (k > 1) ? face_value(scheme, Nz, c, k) : c[k];"),
}


@pytest.fixture(scope="module")
def mutant_libs():
    libs, workdir = mutants.build_many(MUTATIONS, "oracle_npd.c")
    yield libs
    shutil.rmtree(workdir, ignore_errors=True)


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistakes_are_caught(name, mutant_libs):
    L = O.load_library(mutant_libs[name])
    assert failing_pins(L), f"the planted mistake {name!r} passes every pin"
