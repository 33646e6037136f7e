"""Stirred-tank QOIs on the device against the oracle (same seeded inputs):
torque / power number from the IBM reaction torque, scalar CV and mixing
time from the D3Q7 tracer.  Bars: Np and t95 within 0.5% of the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("collision,Q,markers", [("cm", 27, "lattice"), ("trt", 19, "lattice"),
                                                ("cm", 27, "surface"), ("trt", 19, "surface")])
def test_tank_torque_np_and_cv_match_oracle(collision, Q, markers):
    """markers: the reference's blade lattice, or the seeded surface sampler
    of the 256^3 benchmark configuration (one marker per 0.1 lu^2)."""
    tc = cases.tank_case(40, nu=0.02, collision=collision, Q=Q, sigma=3.0, markers=markers)
    steps = 240
    t_o, cv_o, so = cases.run_tank_oracle(tc, steps, prec=4, scalar_D=0.02)
    t_g, cv_g, sg = cases.run_tank_gpu(tc, steps, scalar_D=0.02)
    scale = np.abs(t_o).max()
    assert np.abs(t_g - t_o).max() <= 1e-4 * scale, np.abs(t_g - t_o).max() / scale
    np_o, _ = cases.np_from_torque(t_o, tc, steps)
    np_g, _ = cases.np_from_torque(t_g, tc, steps)
    assert abs(np_g - np_o) <= 5e-3 * abs(np_o)
    np.testing.assert_allclose(cv_g, cv_o, rtol=1e-4)
    t = np.arange(1, steps + 1, dtype=float)
    mo = O.mixing_time(t, cv_o)
    import lbmflow_amd as L

    mg = L.mixing_time_from_cv(t, cv_g)
    assert mg["cv0"] == pytest.approx(mo["cv0"], rel=1e-5)
    assert (mg["t95"] is None) == (mo["t95"] is None)
    if mo["t95"] is not None:
        assert abs(mg["t95"] - mo["t95"]) <= 0.005 * mo["t95"] + 1.0


@pytest.mark.parametrize("lattice", ["hydro", "d3q7"])
def test_compartment_cv_matches_oracle(lattice):
    """compartment_cv on the device against the oracle's restatement
    (qoi.rs:699-739) over the concentration the device holds, as the runner
    evaluates it (runner.rs:1818): counts equal, CV within 1e-12."""
    import lbmflow_amd as L

    tc = cases.tank_case(24, nu=0.02, collision="trt", Q=19, sigma=2.5)
    t_g, cv_g, sg = cases.run_tank_gpu(tc, 30, scalar_D=0.02, d3q7=lattice == "d3q7")
    conc = sg.gather_scalar_concentration(0).astype(np.float64)
    solid = sg.solid()
    for imp_z in (None, tc["T"] / 3.0):  # impeller plane at the clearance C = T/3
        got = sg.compartment_cv(0, imp_z)
        want = O.compartment_cv(sg.dims, conc, solid, imp_z=imp_z)
        for name, cv, cnt in got:
            cv_o, cnt_o = want[name]
            assert cnt == cnt_o, name
            assert (cv is None) == (cv_o is None), name
            if cv is not None:
                assert cv == pytest.approx(cv_o, rel=1e-12), name
    whole = sg.scalar_cv(0)
    assert whole == pytest.approx(O.scalar_cv(conc, (solid == 0).astype(np.uint8)), rel=1e-12)
