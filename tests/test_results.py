"""S-parameter post-processing (host-only entries of libopem.so; no GPU needed): conversions against the oracle and
closed forms, and the files against the reference's own Touchstone output and regression test cases
(tests/golden/sparam: filter_2nd_order.s2p from regression/OpenParEM3D/microstrip/filter_study, the prototype test
cases of regression/OpenParEM3D/WR75/T-Junction and of the lossless partially filled guide)."""
import os

import numpy as np
import pytest

from oracle import results as ores

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "sparam")


@pytest.fixture(scope="module")
def R():
    from openparem_amd import results
    return results


def _rand_s(n, seed):
    rng = np.random.default_rng(seed)
    return 0.4 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)


def test_line_section_impedance_matrix(R):
    """A lossless matched line of electrical length t: S = [[0, e^-jt], [e^-jt, 0]] at Z0 has
    Z11 = -j Z0 cot t, Z21 = -j Z0 csc t (ABCD of a line section)."""
    z0, t = 50.0, 0.7
    S = np.array([[0, np.exp(-1j * t)], [np.exp(-1j * t), 0]])
    Z = R.s2z(S, z0)
    ref = -1j * z0 * np.array([[1 / np.tan(t), 1 / np.sin(t)], [1 / np.sin(t), 1 / np.tan(t)]])
    assert np.abs(Z - ref).max() < 1e-12 * z0
    assert np.abs(R.z2s(Z, z0) - S).max() < 1e-14


def test_one_port_reflection(R):
    """One port: Z = Zo (1 + G)/(1 - G); a matched load (G = 0) of Zo re-referred to Zr reflects (Zo-Zr)/(Zo+Zr)."""
    zo = 35.0 + 4.0j
    g = 0.3 - 0.2j
    assert abs(R.s2z(np.array([[g]]), zo)[0, 0] - zo * (1 + g) / (1 - g)) < 1e-13 * abs(zo)
    assert abs(R.renormalize(np.array([[0.0]]), zo, 50.0)[0, 0] - (zo - 50) / (zo + 50)) < 1e-15


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_conversions_match_oracle(R, n):
    S = _rand_s(n, 10 + n)
    zo = np.linspace(20, 80, n) + 1j * np.linspace(-5, 5, n)
    Z = R.s2z(S, zo)
    assert np.abs(Z - ores.s2z(S, zo)).max() < 1e-12 * np.abs(Z).max()
    assert np.abs(R.z2s(Z, zo) - S).max() < 1e-13
    assert np.abs(R.z2s(Z, zo) - ores.z2s(Z, zo)).max() < 1e-13
    Sr = R.renormalize(S, zo, 50.0)
    assert np.abs(Sr - ores.renormalize(S, zo, 50.0)).max() < 1e-13
    assert np.abs(R.z2s(R.s2z(Sr, 50.0), zo) - S).max() < 1e-13  # back to the ports' own impedances
    assert np.abs(R.renormalize(S, 42.0, 42.0) - S).max() < 1e-14  # same reference: no change


def test_touchstone_reproduces_reference_file(R, tmp_path):
    """Reading the reference's 2-port Touchstone file and writing it again in the same format gives the same
    comment, option and column-name lines byte for byte, the same frequency column, and the same numbers in the same
    order (N11 N21 N12 N22) to the last digit or two that the dB -> magnitude -> dB round trip moves."""
    src = os.path.join(GOLD, "filter_2nd_order.s2p")
    f, S, zref = R.read_touchstone(src)
    assert S.shape == (len(f), 2, 2) and zref == 50.0
    out = tmp_path / "f.s2p"
    R.write_touchstone(out, f, S, fmt="DB", unit="GHz", zref=zref)
    mine, ref = open(out).read().splitlines(), open(src).read().splitlines()
    assert len(mine) == len(ref)
    for a, b in zip(mine, ref):
        if b.startswith(("!", "#")):
            assert a == b
            continue
        ta, tb = a.split(), b.split()
        assert len(ta) == len(tb) == 9 and ta[0] == tb[0]
        for x, y in zip(ta[1:], tb[1:]):
            assert abs(float(x) - float(y)) <= 1e-12 * max(1.0, abs(float(y))) and len(x) <= len(y) + 1


@pytest.mark.parametrize("n", [3, 4, 5])
def test_touchstone_multiport_round_trip(R, tmp_path, n):
    """Three and more ports: row by row, four entries a line, '!row i' on the first line of each row."""
    f = np.array([1e9, 2.5e9])
    S = np.stack([_rand_s(n, 3), _rand_s(n, 4)])
    out = tmp_path / f"x.s{n}p"
    R.write_touchstone(out, f, S, fmt="RI", unit="MHz", zref=50.0, version="2.0.0")
    text = open(out).read().splitlines()
    assert text[0] == "! OpenParEM3D 2.0.0" and text[1] == f"! {n}-port S-parameter data"
    assert f"# MHz S RI R 50" in text
    rows = [ln for ln in text if "!row" in ln]
    assert len(rows) == 2 * n and rows[0].split("!row")[1].strip() == "1"
    f2, S2, _ = R.read_touchstone(out)
    assert np.allclose(f2, f) and np.abs(S2 - S).max() < 1e-14


def test_touchstone_needs_renormalized_data(R, tmp_path):
    from openparem_amd.results import ResultsError
    with pytest.raises(ResultsError, match="renormalized"):
        R.write_touchstone(tmp_path / "a.s2p", [1e9], np.zeros((1, 2, 2)), zref=0.0)


def _cases(path):
    out = []
    for ln in open(path):
        if ln.startswith("#") or not ln.strip():
            continue
        out.append(ln.strip().split(","))
    return out


def _s_from_equal_cases(path, n):
    """S at each frequency of a test-case file whose cases are all 'equal' on dB and degrees."""
    S = {}
    for c in _cases(path):
        f, var, po, pi, val = float(c[1]), c[2], int(c[3]), int(c[4]), float(c[6])
        m = S.setdefault(f, [[[0.0, 0.0] for _ in range(n)] for _ in range(n)])
        m[pi - 1][po - 1][0 if var == "magS(dB)" else 1] = val
    freqs = np.array(sorted(S))
    out = np.array([[[10 ** (e[0] / 20) * np.exp(1j * np.deg2rad(e[1])) for e in row] for row in S[f]]
                    for f in freqs])
    return freqs, out


def test_comparator_passes_reference_cases_on_their_own_values(R, tmp_path):
    """The T-junction's prototype cases (3 ports, 10, 13 and 16 GHz): S rebuilt from their expected values, written as a
    DB-format CSV, passes every case; one entry moved by 1e-9 relative fails exactly the cases that read it."""
    cases = os.path.join(GOLD, "T-Junction_test_cases.csv")
    f, S = _s_from_equal_cases(cases, 3)
    csv = tmp_path / "t.csv"
    R.write_csv(csv, f, S, fmt="DB", unit="GHz", zref=50.0)
    nc, ne, npas = R.process_tests(cases, csv, tmp_path / "report.csv")
    assert nc == len(_cases(cases)) == 54 and ne == nc and npas == nc
    S2 = S.copy()
    S2[1, 2, 0] *= 1 + 1e-9  # entry (portIn 3, portOut 1) at 13 GHz, magnitude only
    R.write_csv(csv, f, S2, fmt="DB", unit="GHz", zref=50.0)
    nc, ne, npas = R.process_tests(cases, csv, tmp_path / "report.csv")
    assert npas == nc - 1
    rep = open(tmp_path / "report.csv").read().splitlines()
    assert rep[1].split(",")[1] == "FAIL" and rep[1].split(",")[3:6] == ["1", "3", "magS(dB)"]


def test_comparator_thresholds_and_wraps(R, tmp_path):
    """The partially filled guide's cases: 'lessthan'/'greaterthan' on dB (|S21| = 0 dB within 1e-5, S11 below
    -100 dB) and 'equal' on a phase, with the phase given 360 degrees away."""
    cases = os.path.join(GOLD, "part_test_cases.csv")
    f = np.array([9e9, 11e9])
    S = np.zeros((2, 2, 2), np.complex128)
    for k, ph in enumerate((84.470987655067, 25.4530954749102)):
        t = np.exp(1j * np.deg2rad(ph - 360))
        S[k] = [[1e-6, t], [t, 1e-6]]
    csv = tmp_path / "p.csv"
    R.write_csv(csv, f, S, fmt="DB", unit="Hz", zref=0.0)
    nc, ne, npas = R.process_tests(cases, csv, tmp_path / "r.csv")
    assert nc == len(_cases(cases)) and npas == nc
    S[0, 0, 0] = 1e-4  # -80 dB breaks the -100 dB bound at 9 GHz
    R.write_csv(csv, f, S, fmt="DB", unit="Hz", zref=0.0)
    assert R.process_tests(cases, csv, tmp_path / "r.csv")[2] == nc - 1


def test_save_as_test_round_trip_and_reference_layout(R, tmp_path):
    """save_as_test writes the reference's prototype-case layout (names <case>_<freq index>_<n>_result, portOut and
    portIn of entry (i, j) as j+1, i+1); its cases for a non-reciprocal S pass on that S and fail on S^T."""
    cases_ref = _cases(os.path.join(GOLD, "T-Junction_test_cases.csv"))
    f, S = _s_from_equal_cases(os.path.join(GOLD, "T-Junction_test_cases.csv"), 3)
    out = tmp_path / "cases.csv"
    R.save_as_test(out, "T-Junction", f, S)
    mine = _cases(out)
    assert [c[:6] for c in mine] == [c[:6] for c in cases_ref]
    for a, b in zip(mine, cases_ref):
        assert abs(float(a[6]) - float(b[6])) <= 1e-12 * abs(float(b[6]))
    Sn = np.stack([_rand_s(3, 7 + k) for k in range(len(f))])
    out2 = tmp_path / "cases2.csv"
    R.save_as_test(out2, "x", f, Sn)
    csv = tmp_path / "n.csv"
    R.write_csv(csv, f, Sn, fmt="DB", unit="GHz", zref=50.0)
    nc, _, npas = R.process_tests(out2, csv, tmp_path / "r.csv")
    assert npas == nc
    R.write_csv(csv, f, np.transpose(Sn, (0, 2, 1)), fmt="DB", unit="GHz", zref=50.0)
    assert R.process_tests(out2, csv, tmp_path / "r.csv")[2] < nc // 2


def test_misuse_and_failures(R, tmp_path):
    """Bad arguments are misuse (-2); a singular conversion and an unwritable path are failures (1), each with a
    message; nothing is written on misuse."""
    from openparem_amd._lib import lib
    from openparem_amd.results import ResultsError
    with pytest.raises(ResultsError, match=r"rc=1\).*singular"):
        R.s2z(np.eye(2), 50.0)  # I - S singular: every port open
    with pytest.raises(ResultsError, match=r"rc=-2\).*format"):
        R.write_csv(tmp_path / "x.csv", [1e9], np.zeros((1, 2, 2)), fmt="XY")
    assert not (tmp_path / "x.csv").exists()
    with pytest.raises(ResultsError, match=r"rc=-2\).*unit"):
        R.write_touchstone(tmp_path / "x.s2p", [1e9], np.zeros((1, 2, 2)), unit="THz")
    with pytest.raises(ResultsError, match=r"rc=1\).*cannot open"):
        R.write_csv(tmp_path / "no" / "such" / "dir.csv", [1e9], np.zeros((1, 2, 2)))
    assert lib().opem_s2z(0, None, None, None) == -2
    with pytest.raises(ResultsError, match="cannot open"):
        R.process_tests(tmp_path / "missing.csv", tmp_path / "missing2.csv", tmp_path / "r.csv")
    bad = tmp_path / "bad_cases.csv"
    bad.write_text("a,9e9,magS(dB),1,1,equal,-3\n")  # 'equal' without a tolerance
    good = tmp_path / "g.csv"
    R.write_csv(good, [9e9], np.zeros((1, 2, 2)) + 0.5)
    with pytest.raises(ResultsError, match="bad test case on line 1"):
        R.process_tests(bad, good, tmp_path / "r.csv")
