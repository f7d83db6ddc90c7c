"""Host side of the drop-in on the CPU: the parameters.xml reader and the
TESTMODE initial loaders of libpdp3_host.so.

The loaders must reproduce the reference's initial state bit for bit: they
are compared with the oracle's restatement (whose output the reference's
smoke goldens pin end to end).  The sampler's inverse-CDF tables are checked
against their defining equations with scipy's incomplete gamma function, and
the Maxwell-Boltzmann branch (T < 0.1 eV, which no golden covers) against the
Gamma(3/2) law it samples.
"""
import os

import numpy as np
import pytest
import scipy.special as sps
import scipy.stats as st

import smoke_case as sc

drv = pytest.importorskip("pdp3_amd.driver")

HERE = os.path.dirname(os.path.abspath(__file__))
TEMPLATE = os.path.join(HERE, "golden", "smoke", "parameters.xml.in")


def write_params(tmp_path, text=None, precision="fp64", result=None):
    if text is None:
        with open(TEMPLATE) as f:
            text = f.read()
    text = text.replace("@RESULT_PATH@", str(result or tmp_path)).replace("@PRECISION@", precision)
    p = tmp_path / "parameters.xml"
    p.write_text(text)
    return str(p)


def test_random_reverse_matches_oracle(orc):
    for power in (3, 4, 5, 7, 9, 11, 13, 15, 23):
        for v in (0, 1, 2, 7, 100, 12345, 999999, 2**31 - 2):
            assert drv.random_reverse(v, power) == orc.random_reverse(v, power)
    # radical inverse closed form: 1 = 0.1 in any base, 5 = 0.21 base 2 ... (base 10 digits reversed)
    assert drv.random_reverse(1, 10) == 0.1
    assert abs(drv.random_reverse(123, 10) - 0.321) < 1e-15
    assert drv.random_reverse(0.99, 7) == 0.0


def test_load_cylindrical_bitwise(orc):
    dr, dz = sc.dr_dz()
    for n1, n2 in ((2e16, 2.05e16), (1e17, 1e17)):
        ours = drv.load_cylindrical(20000, 0.25, 2.0, dr, dz, n1, n2, 0.0)
        ref = orc.load_cylindrical(20000, 0.25, 2.0, dr, dz, n1, n2, 0.0)
        for a, b in zip(ours, ref):
            assert np.array_equal(a, b)


def test_velocity_distribution_bitwise(orc):
    for mass, t in ((sc.EL_MASS, 1.0), (1836 * sc.EL_MASS, 0.1), (sc.EL_MASS, 5.0)):
        ours = drv.velocity_distribution(20000, mass, t)
        ref = orc.velocity_testmode(20000, mass, t)
        for a, b in zip(ours, ref):
            assert np.array_equal(a, b)


def test_bunch_loaders_bitwise(orc):
    dr, dz = sc.dr_dz()
    b = sc.BEAM
    dur = b["length"] / b["velocity"]
    ours = drv.bunch_positions(50000, b["radius"], dr, dz, dur, b["velocity"], b["density"])
    ref = orc.bunch_testmode(50000, b["radius"], dr, dz, dur, b["velocity"], b["density"])
    for a, c in zip(ours, ref):
        assert np.array_equal(a, c)
    u = drv.bunch_inject_draws(5000)
    assert np.array_equal(u, np.array([orc.random_reverse(i, 11) for i in range(5000)]))


def test_tables_solve_their_equations(orc):
    mj = drv.mj_table()
    assert np.array_equal(mj, orc.mj_table())
    y = -30.0 + np.arange(1000) * 42.0 / 999.0
    x = np.exp(mj)
    # -log Q(3, x) = exp(y): scipy's upper function where Q is neither ~1 nor
    # below the FP64 range, its lower function (P = 1 - Q) for small x
    mid = (x > 0.5) & (x < 600)
    lhs = -np.log(sps.gammaincc(3.0, x[mid]))
    assert np.max(np.abs(lhs / np.exp(y[mid]) - 1)) < 1e-9
    lo = x <= 0.5
    assert np.max(np.abs(sps.gammainc(3.0, x[lo]) / -np.expm1(-np.exp(y[lo])) - 1)) < 1e-9
    # above x = 600, Q(3, x) is below the FP64 range: the entries there are
    # covered by the bisection residual of the table solver only
    assert mid.sum() + lo.sum() >= 860

    mb = drv.mb_table()
    y = -19.0 + np.arange(1000) * 21.0 / 999.0
    x = np.exp(mb)
    p = sps.gammainc(1.5, x)
    target = -np.expm1(-np.exp(y))  # P(3/2, x) = 1 - exp(-exp(y))
    assert np.max(np.abs(p / target - 1)) < 1e-9


def test_maxwell_boltzmann_branch_inverts_q():
    # T < 0.1 eV: energy_i = T x_i with Q(3/2, x_i) = U_i, U_i the TESTMODE
    # draw random_reverse(n - i, 9 + i) (src/math/maxwellJuttner.cpp:333-336);
    # x_i from scipy's inverse, up to the table's log-linear interpolation
    n, t = 4000, 0.05
    e = drv.maxwell_juttner(n, t) / t
    u = np.array([drv.random_reverse(n - i, 9 + i) for i in range(n)])
    lnln = np.log(-np.log(u))
    inside = (lnln > -19) & (lnln < 2)
    x = sps.gammainccinv(1.5, u[inside])
    assert inside.sum() > 0.9 * n
    assert np.max(np.abs(e[inside] / x - 1)) < 2e-4
    # outside the table the reference's own closed forms apply
    assert np.all(e[lnln <= -19] == 1.0)


def test_maxwell_juttner_branch_replay():
    # T >= 0.1: rejection sampling of gamma (src/math/maxwellJuttner.cpp:350-389)
    # replayed with scipy: x solves Q(3, x) = (1 - U) Q(3, 1/T) for the draw
    # U = random_reverse(n + i - k, 13), gamma = T x is accepted when
    # random_reverse(n - i + k, 15) < sqrt(1 - 1/gamma^2)
    n = 3000
    for t in (0.1, 1.0):
        e = drv.maxwell_juttner(n, t)
        q0 = sps.gammaincc(3.0, 1.0 / t)
        checked = 0
        for i in range(n):
            k = 0
            while True:
                u = drv.random_reverse(n + i - k, 13)
                gamma = t * sps.gammainccinv(3.0, (1 - u) * q0)
                a = drv.random_reverse(n - i + k, 15)
                beta = np.sqrt(1 - 1 / gamma ** 2)
                k += 1
                if abs(a - beta) < 1e-3:  # decision within the table's interpolation error
                    gamma = None
                    break
                if a < beta:
                    break
            if gamma is None:
                continue
            assert abs((e[i] + 1) / gamma - 1) < 2e-4, (t, i, e[i] + 1, gamma)
            checked += 1
        assert checked > 0.95 * n


def test_parameters_reader(tmp_path):
    p = drv.parameters(write_params(tmp_path))
    assert p["n_grid_r"] == "255" and p["n_grid_z"] == "2047"
    assert float(p["r_size"]) == 0.25 and float(p["z_size"]) == 2.0
    assert float(p["end_time"]) == 2.2e-11 and float(p["step_interval"]) == 1e-12
    assert p["species0"].split(",")[:3] == ["Electrons", "-1", "1"]
    assert float(p["species0"].split(",")[3]) == 1e5
    assert p["species1"].split(",")[:3] == ["Ions", "1", "1836"]
    beam = p["beam"].split(",")
    assert beam[0] == "Beam" and beam[1] == "-1" and float(beam[5]) == 2e8 and beam[6] == "1000000"
    assert p["probe0"] == "0,E_r,none,0,0,32,256,5"
    assert p["probe10"] == "0,T,Electrons,100,100,150,150,5"
    assert p["probe12"] == f"1,E_z,none,0,100,255,-1,5"      # col: r over the whole grid
    assert p["probe13"] == f"2,E_r,none,10,0,-1,2047,5"      # row
    assert p["probe14"] == "3,rho_beam,none,10,50,-1,-1,5"   # dot
    assert p["precision"] == "0" and p["pusher"] == "0" and p["sort_interval"] == "0"
    assert p["debug"] == "0" and p["compress"] == "0"


def test_parameters_debug_and_device(tmp_path):
    text = open(TEMPLATE).read().replace("<debug>false</debug>", "<debug>TRUE</debug>")
    text = text.replace("<pusher>boris</pusher>", "<pusher>vay</pusher>")
    p = drv.parameters(write_params(tmp_path, text, precision="fp32"))
    assert p["debug"] == "1" and p["n_grid_r"] == "15" and p["n_grid_z"] == "127"
    assert p["precision"] == "1" and p["pusher"] == "1"
    assert p["data_root"] == "pdp3_result_debug"


@pytest.mark.parametrize("edit,code", [
    (("<n_grid_r>255</n_grid_r>", ""), 1),                     # missing child
    (('type="dot"', 'type="blob"'), 1),                         # unknown probe type
    (('component="T" specie="Ions" />', 'specie="Ions" />'), 1),  # missing attribute
    (("<precision>@PRECISION@</precision>", "<precision>fp16</precision>"), 1),
    (("</initial_parameters>", ""), 78),                        # not well-formed
])
def test_parameters_errors(tmp_path, edit, code):
    text = open(TEMPLATE).read().replace(*edit)
    with pytest.raises(drv.DriverError) as ei:
        drv.parameters(write_params(tmp_path, text))
    assert ei.value.code == code


def test_parameters_missing_file(tmp_path):
    with pytest.raises(drv.DriverError) as ei:
        drv.parameters(str(tmp_path / "nope.xml"))
    assert ei.value.code == 78


def test_xml_entities_comments(tmp_path):
    text = open(TEMPLATE).read().replace("<particle_beam name=\"Beam\">",
                                         "<!-- a <comment> --><particle_beam name=\"B&amp;m&#65;\">")
    text = text.replace("<data_root>pdp3_result</data_root>", "<data_root>  res\n   &lt;x&gt;  </data_root>")
    p = drv.parameters(write_params(tmp_path, text))
    assert p["beam"].split(",")[0] == "B&mA"
    assert p["data_root"] == "res <x>"
