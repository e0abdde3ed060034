"""Pins the CPU oracle against the reference build's own outputs.

The fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py
from the reference setup.out/core.out) hold record 0, the SHA-256 digest of
E, H and the marker array of every one of 20 later records, the reference
diag.dat and the MC_GAUSS_TEST build's gauss_test.dat.  The oracle must
reproduce every record bit for bit, marker order included (so the J
summation order is the reference's too), and the printed diagnostics to
their 7 printed digits.
"""
import numpy as np
import pytest

from oracle_lib import (field_digest, load_fixture, marker_digest,
                        oracle_from_fixture)

CASES = ["periodic", "laser_mur", "mixed", "spot", "tfsf", "tfsf_back", "vsp"]


@pytest.mark.parametrize("case", CASES)
def test_bit_exact_records(case):
    fx = load_fixture(case)
    o = oracle_from_fixture(fx)
    for r in range(1, len(fx["digest_E"]) + 1):
        o.step()
        E, H = o.fields()
        M = o.markers()
        assert field_digest(E) == fx["digest_E"][r - 1], (case, r, "E")
        assert field_digest(H) == fx["digest_H"][r - 1], (case, r, "H")
        assert marker_digest(M) == fx["digest_M"][r - 1], (case, r, "markers")
        # sys_%06d.txt prints the time with %.14e
        assert abs(o.time - fx["times"][r - 1]) <= 1e-13 * abs(o.time)
        if r == 1:
            np.testing.assert_array_equal(E, fx["E1"])
            np.testing.assert_array_equal(H, fx["H1"])


@pytest.mark.parametrize("case", CASES)
def test_diagnostics_match_reference_files(case):
    fx = load_fixture(case)
    o = oracle_from_fixture(fx)
    rows, gs = [], []
    for _ in range(len(fx["digest_E"])):
        d, g = o.step(gauss=True)
        rows.append(d)
        gs.append(g)
    rows = np.array(rows)[1:]            # diag.dat skips the t = 0 point
    ref = fx["diag"][:, :6]
    # diag.dat is printed with %e (7 significant digits)
    np.testing.assert_allclose(rows, ref, rtol=6e-7, atol=0)
    g = fx["gauss"][:, 1:]
    np.testing.assert_allclose(np.array(gs), g, rtol=6e-7, atol=0)


def _tfsf_first_mismatch(fx, dataE, dataH, mode=None, nframes=None):
    """First record (1-based) whose E or H digest differs from the
    reference's when the oracle plays the given face data, or 0."""
    from oracle_lib import Oracle, tfsf_arrays
    reg, m, nf, _, _ = tfsf_arrays(fx)
    o = Oracle(fx["n"], fx["h"], fx["L"], fx["tau"][0], fx["bc"], 0.0)
    o.set_fields(fx["E0"], fx["H0"])
    o.set_markers(fx["M0"])
    if dataE is not None:
        o.add_tfsf(reg, m if mode is None else mode, nf if nframes is None else nframes,
                   dataE, dataH)
    o.init()
    for r in range(1, len(fx["digest_E"]) + 1):
        o.step()
        E, H = o.fields()
        if field_digest(E) != fx["digest_E"][r - 1] or field_digest(H) != fx["digest_H"][r - 1]:
            return r
    return 0


@pytest.mark.parametrize("case", ["tfsf", "tfsf_back"])
def test_tfsf_fixture_pins_the_interface(case):
    """The TFSF fixtures are sensitive to each plausible mistake: no
    interface, E and H recordings swapped, the play direction reversed, and
    an interface that stays live one frame longer (forward case: the cap
    call that switches it off at the start of a step)."""
    from oracle_lib import tfsf_arrays
    fx = load_fixture(case)
    _, mode, nf, dataE, dataH = tfsf_arrays(fx)
    assert _tfsf_first_mismatch(fx, dataE, dataH) == 0
    assert _tfsf_first_mismatch(fx, None, None) == 1
    assert _tfsf_first_mismatch(fx, dataH, dataE) == 1
    assert _tfsf_first_mismatch(fx, dataE, dataH, mode=2 - mode) == 1
    if case == "tfsf":
        # one frame more keeps the interface live through step 14 (F = 13)
        ext = [None if x is None else np.concatenate([x, x[-x.size // nf:]]) for x in dataE]
        exth = [None if x is None else np.concatenate([x, x[-x.size // nf:]]) for x in dataH]
        assert _tfsf_first_mismatch(fx, ext, exth, nframes=nf + 1) == 14


def test_vsp_fixture_pins_the_convolution():
    """The VSP fixture holds a run with the (beta, alpha, beta) convolution at
    alpha = 0.8: without the convolution, or with the weight mistaken, the
    oracle leaves the reference's records at the first step; the sweep is
    the identity at alpha = 1 and keeps sums (2 beta + alpha = 1) along
    periodic lines."""
    from oracle_lib import Oracle
    fx = load_fixture("vsp")

    def first_mismatch(alpha):
        o = Oracle(fx["n"], fx["h"], fx["L"], fx["tau"][0], fx["bc"], 0.0)
        o.set_fields(fx["E0"], fx["H0"])
        o.set_markers(fx["M0"])
        if alpha is not None:
            o.set_vsp(alpha)
        o.init()
        for r in range(1, len(fx["digest_E"]) + 1):
            o.step()
            E, _ = o.fields()
            if field_digest(E) != fx["digest_E"][r - 1]:
                return r
        return 0
    assert first_mismatch(0.8) == 0
    assert first_mismatch(None) == 1
    assert first_mismatch(0.75) == 1
    # the same box without the option: the "mixed" fixture's deck and record 0
    mx = load_fixture("mixed")
    assert np.array_equal(mx["M0"]["x"], fx["M0"]["x"])
