"""ScaleNegativeTracers (src/Utils/negative_tracers.jl) in the oracle: the
reference's own test (test/test_utils.jl:7-22: N = 2, P = -1 becomes N = 1,
P = 0), the conservation the scheme exists for (the scaled group total
equals the original total when it is not negative), the fill value, NaN
propagation, the host's conserved groups (coupling_utils.jl) and the
scaling inside a stepped column.  CPU only."""
import math

import numpy as np
import pytest

import oracle as O
from oceanbiome import LOBSTER, NPZD, CarbonateSystem, RectilinearGrid, abi

NAN = float("nan")


def _npzd_scaling(fill=NAN):
    return abi.negative_scaling([(["P", "Z", "N", "D"], [1, 1, 1, 1], fill)], ("N", "P", "Z", "T", "D"))


def test_reference_case():
    """test_utils.jl:7-22 (set!(model, N = 2, P = -1) then one tiny step)."""
    out = O.scale_negatives(_npzd_scaling(), [2.0, -1.0, 0.0, 10.0, 0.0])
    assert out[0] == pytest.approx(1.0) and out[1] == 0.0
    assert out[3] == 10.0                    # T is not a member


def test_group_total_is_conserved():
    rng = np.random.default_rng(4)
    s = abi.negative_scaling([(["a", "b", "c", "d"], [1.0, 2.0, 0.5, 6.56], NAN)], ("a", "b", "c", "d", "e"))
    for _ in range(200):
        v = rng.normal(1.0, 1.0, 5)
        total = v[0] + 2 * v[1] + 0.5 * v[2] + 6.56 * v[3]
        out = O.scale_negatives(s, v)
        if total >= 0:
            assert np.all(out[:4] >= 0)
            new_total = out[0] + 2 * out[1] + 0.5 * out[2] + 6.56 * out[3]
            assert new_total == pytest.approx(total, rel=1e-13, abs=1e-14)
            # members keep their proportions
            pos = v[:4] > 0
            if pos.sum() >= 2:
                r = out[:4][pos] / v[:4][pos]
                assert np.allclose(r, r[0], rtol=1e-14)
        assert out[4] == v[4]


def test_negative_total_uses_the_fill_value():
    s = _npzd_scaling(fill=0.0)
    out = O.scale_negatives(s, [-3.0, 1.0, 0.5, 10.0, 0.0])
    assert list(out) == [0.0, 0.0, 0.0, 10.0, 0.0]
    out = O.scale_negatives(_npzd_scaling(), [-3.0, 1.0, 0.5, 10.0, 0.0])
    assert out[0] == 0.0 and math.isnan(out[1]) and math.isnan(out[2])


def test_nan_member_spreads_to_positive_members():
    out = O.scale_negatives(_npzd_scaling(), [NAN, 1.0, -0.1, 10.0, 0.5])
    assert math.isnan(out[0]) and math.isnan(out[1]) and out[2] == 0.0 and math.isnan(out[4])


def test_host_conserved_groups_follow_the_reference():
    grid = RectilinearGrid(size=(2, 2, 4), extent=(1, 1, 4))
    m = NPZD(grid, scale_negatives=True)
    assert m.conserved_tracers() == [(["P", "Z", "N", "D"], [1.0] * 4)]
    m = LOBSTER(grid, carbonate_system=CarbonateSystem(), scale_negatives=True)
    g = m.conserved_tracers()
    assert g[0][0] == ["P", "Z", "NO₃", "NH₄", "sPOM", "bPOM", "DOM"]
    R, rho = 6.56, 0.1
    assert g[1][0] == ["P", "Z", "DIC", "sPOM", "bPOM", "DOM"]
    assert g[1][1] == pytest.approx([(1 + rho) * R, R, 1.0, R, R, R])
    s = m.config(abi.OBME_FP64).scaling
    names = m.required_biogeochemical_tracers()
    assert s.n_groups == 2 and s.n_members[0] == 7 and s.n_members[1] == 6
    assert [names[s.member[1][k]] for k in range(6)] == g[1][0]
    assert math.isnan(s.invalid_fill_value[0])
    assert NPZD(grid).config(abi.OBME_FP64).scaling.n_groups == 0


def test_scaling_in_a_stepped_column():
    """A column with a negative plankton value: after one step with the
    modifier the first tendency evaluation sees the scaled state, and the
    column's nitrogen is the scaled total's evolution."""
    grid = RectilinearGrid(size=(1, 1, 6), extent=(1, 1, 30), topology=("Periodic", "Periodic", "Bounded"))
    base = NPZD(grid)
    scaled = NPZD(grid, scale_negatives=True)
    fields = []
    for v in (5.0, -0.2, 0.1, 12.0, 0.3):
        f = grid.new_field()
        grid.interior(f)[...] = v
        fields.append(f)
    a = [f.copy() for f in fields]
    O.grid_run(scaled.config(abi.OBME_FP64), a, 1, 1e-10)
    # with a negligible step the state is the scaled initial state
    b = [f.copy() for f in fields]
    O.grid_scale_negatives(scaled.config(abi.OBME_FP64), b)
    for x, y in zip(a, b):
        assert np.allclose(grid.interior(x), grid.interior(y), rtol=1e-9, atol=1e-12)
    assert np.all(grid.interior(b[1]) == 0.0)
    c = [f.copy() for f in fields]
    O.grid_run(base.config(abi.OBME_FP64), c, 1, 1e-10)
    # without the modifier the negative P reaches the PAR power (P R/r)^e of
    # a negative base: NaN (a DomainError on the reference's CPU path)
    assert np.all(np.isnan(grid.interior(c[1])))
