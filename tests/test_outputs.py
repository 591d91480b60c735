"""Text outputs byte for byte as the reference's std::println writes them (port_sweep.txt,
polar_<n>.txt: src/emw_ana_antenna.cpp:150-203; bistatic_rcs_<n>.txt:
src/emw_ana_bistatic_rcs.cpp:97-117).  std::format's "{}" of a double is std::to_chars' shortest
form: the fewest characters that read back to the value, in fixed or scientific notation, fixed on
a tie, and among equally short strings the one closest to the value (so a large integer-valued
double prints as its exact integer).  The library calls std::to_chars; the tests restate the
rule independently from Python's repr (also a shortest round-trip algorithm)."""
import math

import numpy as np
import pytest

import frico_amd as fa


def shortest(v: float) -> str:
    """The "{}" rule restated: shortest digits from repr, then the shorter of %f / %e forms."""
    v = float(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    if math.isinf(v):
        return "-inf" if v < 0 else "inf"
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    r = repr(abs(v))
    if "e" in r:
        mant, ex = r.split("e")
        ex = int(ex)
    else:
        mant, ex = r, 0
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    digits = (ip + fp).lstrip("0")
    # decimal exponent of the first significant digit
    lead = len(ip.lstrip("0"))
    point = (lead if lead else -(len(fp) - len(fp.lstrip("0")))) + ex
    digits = digits.rstrip("0") or "0"
    if digits == "0":
        return sign + "0"
    e10 = point - 1
    sci = digits[0] + ("." + digits[1:] if len(digits) > 1 else "") + "e" + ("-" if e10 < 0 else "+") + f"{abs(e10):02d}"
    if point <= 0:
        fix = "0." + "0" * (-point) + digits
    elif point >= len(digits):
        # all zero-padded forms are equally short; the one closest to the value is the exact integer
        fix = str(int(abs(v)))
    else:
        fix = digits[:point] + "." + digits[point:]
    return sign + (fix if len(fix) <= len(sci) else sci)


@pytest.mark.parametrize("v,s", [(6943004927317386240.0, "6943004927317386240"), (3e8, "3e+08"), (2.5e8, "2.5e+08"), (123456789.0, "123456789"), (0.1, "0.1"),
                                 (72.203395, "72.203395"), (1e-5, "1e-05"), (1e22, "1e+22"), (100.0, "100"),
                                 (10000.0, "10000"), (100000.0, "1e+05"), (0.001, "0.001"), (0.0001, "1e-04"),
                                 (1.5e-7, "1.5e-07"), (-0.0, "-0"), (0.0, "0"), (float("inf"), "inf"),
                                 (float("-inf"), "-inf"), (50.0, "50"), (-121.852668, "-121.852668")])
def test_known_values(v, s):
    assert fa.format_double(v) == s
    assert shortest(v) == s


def test_random_values_follow_the_rule():
    rng = np.random.default_rng(7)
    vals = np.concatenate([rng.uniform(-1, 1, 5000) * 10.0 ** rng.integers(-30, 30, 5000),
                           rng.integers(-10 ** 9, 10 ** 9, 5000).astype(np.float64),
                           np.array([round(x, int(d)) for x, d in zip(rng.uniform(-1000, 1000, 5000),
                                                                         rng.integers(0, 8, 5000))]),
                           rng.standard_normal(5000).view(np.float64)])
    for v in vals:
        s = fa.format_double(float(v))
        assert s == shortest(float(v)), (v, s)
        assert float(s) == v


def test_port_sweep_file(tmp_path):
    f = [2.5e8, 2.6e8, 2.7e8]
    z = [38.767517 - 121.852668j, 43.895262 - 97.077209j, 49.688086 - 72.641564j]
    swr = [9.621, 6.148, 3.865]
    p = tmp_path / "port_sweep.txt"
    fa.write_port_sweep(p, 50.0, f, z, swr)
    want = "# Port sweep results, Z0 = 50.0 Ohm\n# step  frequency    Re(Z)    Im(Z)    SWR\n"
    for i in range(3):
        want += f"{i} {shortest(f[i])} {shortest(z[i].real)} {shortest(z[i].imag)} {shortest(swr[i])}\n"
    assert p.read_bytes() == want.encode()
    assert p.read_text().splitlines()[2] == "0 2.5e+08 38.767517 -121.852668 9.621"


def test_radiation_diagram_file(tmp_path):
    rng = np.random.default_rng(1)
    g = rng.uniform(0, 2, (3, 360))
    g[1, 7] = 0.0  # a null prints -inf in dB, as log10(0) does in the reference
    p = tmp_path / "polar_0.txt"
    fa.write_radiation_diagram(p, 3e8, g)
    lines = p.read_text().split("\n")
    assert lines[0] == "# FRICO radiation patterns. Frequency = 3e+08"
    assert lines[1] == "# angle XY_lin XY_dB YZ_lin YZ_dB XZ_lin XZ_dB"
    assert len(lines) == 363 and lines[-1] == ""
    for i in (0, 7, 359):
        cols = [str(i)]
        for pl in range(3):
            # math.log10 is the C library's log10, as std::log10 in the reference
            cols += [shortest(g[pl, i]), shortest(10.0 * math.log10(g[pl, i]) if g[pl, i] > 0 else -math.inf)]
        assert lines[2 + i] == " ".join(cols)
    assert lines[9].split()[4] == "-inf"


def test_bistatic_rcs_file(tmp_path):
    t = np.arange(359 * 4, dtype=np.float64).reshape(359, 4) * 0.37
    p = tmp_path / "bistatic_rcs_0.txt"
    fa.write_bistatic_rcs(p, t)
    lines = p.read_text().splitlines()
    assert lines[:2] == ["# FRICO bistatic RCS computation", "# Angle, VV_theta, HH_theta, VV_phi, HH_phi"]
    assert len(lines) == 361
    assert lines[2 + 100] == "100 " + " ".join(shortest(x) for x in t[100])


def test_write_errors(tmp_path):
    with pytest.raises(fa.FricoError):
        fa.write_bistatic_rcs(tmp_path / "missing_dir" / "x.txt", np.zeros((359, 4)))


def test_sweep_frequencies_accumulate_like_init_sweep():
    """init_sweep (src/emw_solver.cpp:460-471) accumulates freq += step while freq <= end.  With a
    step of 0.1 from 1 to 2 the eleventh value is 2.0000000000000004 > 2, so the sweep has ten
    points, not eleven; with a step that is a power of two the end point stays."""
    import frico_amd
    f = frico_amd.sweep_frequencies(1.0, 2.0, 0.1)
    assert f.size == 10
    assert f[0] == 1.0 and f[1] == 1.1 and f[2] == 1.2000000000000002
    assert 1.0 + 10 * 0.1 == 2.0 and f[-1] < 2.0 - 0.05  # a multiplied sweep would keep 2.0
    g = frico_amd.sweep_frequencies(100e6, 500e6, 100e6)
    assert g.tolist() == [100e6, 200e6, 300e6, 400e6, 500e6]
    h = frico_amd.sweep_frequencies(1.0, 2.0, 0.25)
    assert h.tolist() == [1.0, 1.25, 1.5, 1.75, 2.0]
    assert frico_amd.sweep_frequencies(3.0, 2.0, 0.5).size == 0
    with pytest.raises(frico_amd.FricoError):
        frico_amd.sweep_frequencies(1.0, 2.0, 0.0)
