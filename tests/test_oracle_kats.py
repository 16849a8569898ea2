"""The oracle against the known-answer values of the reference's own tests
(fixtures in tests/golden/reference_kats.json, see make_golden.py)."""
import math

import numpy as np
import pytest


def test_permanent_pair_coulomb_and_ewald(oracle, kats):
    w = kats["water_dimer"]
    tol = w["tolerance"]
    for key, beta in (("perm_coulomb", 0.0), ("perm_ewald", w["beta"])):
        e, gi, ti, tk = oracle.perm_pair(w["r"], w["Qi"], w["Qk"], beta)
        assert abs(e - w[key]["energy"]) < tol
        assert np.allclose(ti, w[key]["torque_i"], atol=tol, rtol=0)
        assert np.allclose(tk, w[key]["torque_k"], atol=tol, rtol=0)


def test_permanent_pair_gradient_is_energy_derivative(oracle, kats):
    w = kats["water_dimer"]
    r = np.array(w["r"])
    e, gi, _, _ = oracle.perm_pair(r, w["Qi"], w["Qk"], w["beta"])
    h = 1e-6
    for d in range(3):
        rp, rm = r.copy(), r.copy()
        rp[d] += h
        rm[d] -= h
        fd = (oracle.perm_pair(rp, w["Qi"], w["Qk"], w["beta"])[0] - oracle.perm_pair(rm, w["Qi"], w["Qk"], w["beta"])[0]) / (2 * h)
        # r = x_k - x_i, so dE/dx_i = -dE/dr
        assert abs(gi[d] + fd) < 1e-8


def test_polarization_pair_coulomb_ewald(oracle, kats):
    w = kats["water_dimer"]
    tol = w["tolerance"]
    for key, beta, ui, uk in (("pol_coulomb", 0.0, w["Ui"], w["Uk"]), ("pol_ewald", w["beta"], w["UiEwald"], w["UkEwald"])):
        e, gi, ti, tk = oracle.pol_pair(w["r"], w["Qi"], w["Qk"], ui, ui, uk, uk, beta)
        assert abs(e - w[key]["energy"]) < tol
        assert np.allclose(gi, w[key]["grad_i"], atol=tol, rtol=0)
        assert np.allclose(ti, w[key]["torque_i"], atol=tol, rtol=0)
        assert np.allclose(tk, w[key]["torque_k"], atol=tol, rtol=0)


def test_polarization_thole_correction(oracle, kats):
    """The reference quotes the Thole correction subtracted from the bare
    Coulomb term: damped - undamped = -correction."""
    w = kats["water_dimer"]
    tol = w["tolerance"]
    pd = 1.0 / math.sqrt(w["AiAk"])
    e0, g0, ti0, tk0 = oracle.pol_pair(w["r"], w["Qi"], w["Qk"], w["Ui"], w["Ui"], w["Uk"], w["Uk"], 0.0)
    e1, g1, ti1, tk1 = oracle.pol_pair(w["r"], w["Qi"], w["Qk"], w["Ui"], w["Ui"], w["Uk"], w["Uk"], 0.0,
                                       pd, pd, w["thole"], w["thole"])
    ref = w["pol_thole_correction"]
    assert abs((e0 - e1) - ref["energy"]) < tol
    assert np.allclose(g0 - g1, ref["grad_i"], atol=tol, rtol=0)
    assert np.allclose(ti0 - ti1, ref["torque_i"], atol=tol, rtol=0)
    assert np.allclose(tk0 - tk1, ref["torque_k"], atol=tol, rtol=0)


def test_wrong_separation_sign_fails(oracle, kats):
    """The pinning is sensitive to the r = x_k - x_i convention."""
    w = kats["water_dimer"]
    e, _, _, _ = oracle.perm_pair(-np.array(w["r"]), w["Qi"], w["Qk"], w["beta"])
    assert abs(e - w["perm_ewald"]["energy"]) > 1e-6


@pytest.mark.parametrize("case", range(7))
def test_bspline(oracle, kats, case):
    c = kats["bspline"]["cases"][case]
    th = oracle.bspline(c["x"], c["order"], c["order"] - 1)
    tol = kats["bspline"]["tolerance"]
    assert np.allclose(th[:, 0], c["c"], atol=tol, rtol=0)
    assert np.allclose(th[:, 1], c["d"], atol=tol, rtol=0)


def test_bspline_partition_of_unity_and_derivative_sums(oracle):
    for x in np.linspace(0, 1, 11):
        th = oracle.bspline(x, 5, 3)
        assert abs(th[:, 0].sum() - 1.0) < 1e-14
        for j in (1, 2, 3):
            assert abs(th[:, j].sum()) < 1e-13


def test_bspline_derivatives_by_finite_difference(oracle):
    h = 1e-6
    for x in (0.1, 0.37, 0.8):
        th = oracle.bspline(x, 5, 3)
        tp, tm = oracle.bspline(x + h, 5, 3), oracle.bspline(x - h, 5, 3)
        for j in (1, 2, 3):
            assert np.allclose(th[:, j], (tp[:, j - 1] - tm[:, j - 1]) / (2 * h), atol=1e-7)


def test_erf_known_answers(kats):
    # the oracle calls libm erfc; these are the reference's ErfTest values
    for x, v in kats["erf"]["cases"]:
        assert abs(math.erf(x) - v) <= kats["erf"]["tolerance"]


@pytest.mark.parametrize("dims", [(32, 32, 32), (32, 45, 21), (24, 24, 24), (20, 18, 10)])
def test_convolution_identity(oracle, dims):
    """Complex3DParallelTest: convolution with recip = 1 returns N * input."""
    nx, ny, nz = dims
    rng = np.random.default_rng(3)
    data = np.zeros(2 * nx * ny * nz)
    data[0::2] = rng.random(nx * ny * nz)
    out = oracle.convolution(data, np.ones(nx * ny * nz), nx, ny, nz)
    assert np.allclose(out[0::2] / (nx * ny * nz), data[0::2], atol=1e-14, rtol=0)


def test_fft_matches_numpy(oracle):
    nx, ny, nz = 12, 10, 9
    rng = np.random.default_rng(5)
    z = rng.random((nz, ny, nx)) + 1j * rng.random((nz, ny, nx))
    data = np.empty(2 * z.size)
    data[0::2] = z.real.ravel()
    data[1::2] = z.imag.ravel()
    out = oracle.fft3d(data, nx, ny, nz, False)
    ref = np.fft.fftn(z)
    assert np.allclose(out[0::2], ref.real.ravel(), atol=1e-11)
    assert np.allclose(out[1::2], ref.imag.ravel(), atol=1e-11)
