"""CPU checks of the device flux functions of the learned hybrid.

tests/hostflux.cpp builds the __host__ __device__ flux functions of
flux.hpp and indicator.hpp for the host (hipcc's host pass, no GPU needed).
They are pinned here against the CPU oracle's WENO5 face flux and against
what the mathematics fixes:

* the linear fifth-order combination (2, -13, 47, 27, -3) / 60 and its
  mirror reproduce the face value of cell averages of every polynomial of
  degree <= 4 (five moment conditions fix the five coefficients);
* on linear-ramp stencils every smoothness indicator of WENO5 is equal, so
  WENO5 (weno5.c:35-62, convective_flux.c:48-124) is the linear scheme and
  the linear flux must equal the oracle's face flux to round-off, for the
  Lax-Friedrichs and Steger-Warming splits on the mean and Roe states;
* the closed form of the Lax-Friedrichs linear flux, A (U+ + U-) / 2 +
  lambda* (U+ - U-) / 2, equals its characteristic form on any stencil;
* the fast-arithmetic WENO5 flux of the plain sweeps agrees with the
  oracle's to round-off on random stencils.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

from test_oracle import _face_flux

VARIANTS = [(0, 0), (1, 0), (0, 1), (1, 1)]     # (Roe average, Steger-Warming split)


@pytest.fixture(scope="module")
def hostflux(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = str(tmp_path_factory.mktemp("hostflux") / "hostflux.so")
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-fno-fast-math", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "artracfd_amd", "csrc", "device"),
                    os.path.join(ROOT, "tests", "hostflux.cpp"), "-o", out], check=True, capture_output=True)
    lib = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.host_linear_flux.argtypes = [ctypes.c_int] * 4 + [dp, dp]
    lib.host_weno_flux.argtypes = [ctypes.c_int] * 3 + [dp, dp]
    lib.host_linear_sums.argtypes = [dp, dp, dp]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def linear_flux(lib, st, s, roe, sw, closed):
    st = np.ascontiguousarray(st, dtype=np.float64)
    F = np.zeros(5)
    lib.host_linear_flux(s, roe, sw, closed, _p(st), _p(F))
    return F


def weno_flux(lib, st, s, roe, sw):
    st = np.ascontiguousarray(st, dtype=np.float64)
    F = np.zeros(5)
    lib.host_weno_flux(s, roe, sw, _p(st), _p(F))
    return F


def _state(rho, vel, p):
    vel = np.asarray(vel, dtype=np.float64)
    return np.array([rho, *(rho * vel), p / 0.4 + 0.5 * rho * vel @ vel])


def _random_stencil(rng, ramp):
    """Six conservative states: a linear ramp U_a + m dU, or independent
    random states."""
    if ramp:
        base = _state(rng.uniform(0.8, 1.5), rng.uniform(-0.8, 0.8, 3), rng.uniform(0.8, 1.5))
        dU = base * rng.uniform(-0.03, 0.03, 5)
        return np.array([base + (m - 2.5) * dU for m in range(6)])
    return np.array([_state(rng.uniform(0.5, 2.0), rng.uniform(-1, 1, 3), rng.uniform(0.5, 2.0)) for _ in range(6)])


def test_linear_combination_reproduces_quartic_cell_averages(hostflux):
    """Cells of unit width centred at m = -2..3 around the face at 1/2: the
    upwind (nodes -2..2) and mirrored (3..-1) combinations of the cell
    averages of x^k, k <= 4, are both the point value at the face, so their
    sum is twice it and their difference zero."""
    for k in range(5):
        avg = np.array([((m + 0.5) ** (k + 1) - (m - 0.5) ** (k + 1)) / (k + 1) for m in range(-2, 4)])
        st = np.ascontiguousarray(np.repeat(avg[:, None], 5, axis=1))
        Sg, Dl = np.zeros(5), np.zeros(5)
        hostflux.host_linear_sums(_p(st), _p(Sg), _p(Dl))
        np.testing.assert_allclose(Sg, 2 * 0.5 ** k, rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(Dl, 0.0, atol=1e-14)
    # degrees 5 and 6 are not reproduced (their errors show in the
    # difference and in the sum): the test above is not vacuous
    for k, which in ((5, "diff"), (6, "sum")):
        avg = np.array([((m + 0.5) ** (k + 1) - (m - 0.5) ** (k + 1)) / (k + 1) for m in range(-2, 4)])
        st = np.ascontiguousarray(np.repeat(avg[:, None], 5, axis=1))
        Sg, Dl = np.zeros(5), np.zeros(5)
        hostflux.host_linear_sums(_p(st), _p(Sg), _p(Dl))
        err = abs(Dl[0]) if which == "diff" else abs(Sg[0] - 2 * 0.5 ** k)
        assert err > 1e-2, (k, err)


@pytest.mark.parametrize("s", [0, 1, 2])
@pytest.mark.parametrize("roe, sw", VARIANTS)
def test_linear_flux_equals_oracle_weno5_on_linear_ramps(hostflux, oracle_mod, s, roe, sw):
    rng = np.random.default_rng(100 * s + 10 * roe + sw)
    for _ in range(20):
        st = _random_stencil(rng, ramp=True)
        ref = _face_flux(oracle_mod, list(st), s, 1, roe, sw)
        scale = np.abs(ref).max()
        forms = [0] if sw else [0, 1]          # the closed form exists for the Lax-Friedrichs split
        for closed in forms:
            F = linear_flux(hostflux, st, s, roe, sw, closed)
            assert np.abs(F - ref).max() <= 1e-13 * scale, (closed, F, ref)


@pytest.mark.parametrize("s", [0, 1, 2])
@pytest.mark.parametrize("roe", [0, 1])
def test_closed_form_llf_equals_characteristic_form(hostflux, s, roe):
    rng = np.random.default_rng(7 + s + 3 * roe)
    for _ in range(50):
        st = _random_stencil(rng, ramp=False)
        a = linear_flux(hostflux, st, s, roe, 0, 0)
        b = linear_flux(hostflux, st, s, roe, 0, 1)
        assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("s", [0, 1, 2])
@pytest.mark.parametrize("roe, sw", VARIANTS)
def test_fast_weno5_flux_matches_oracle(hostflux, oracle_mod, s, roe, sw):
    rng = np.random.default_rng(31 + 100 * s + 10 * roe + sw)
    for ramp in (False, True):
        for _ in range(30):
            st = _random_stencil(rng, ramp)
            ref = _face_flux(oracle_mod, list(st), s, 1, roe, sw)
            F = weno_flux(hostflux, st, s, roe, sw)
            assert np.abs(F - ref).max() <= 1e-12 * np.abs(ref).max(), (F, ref)


def test_linear_flux_differs_from_weno5_where_the_data_are_not_smooth(hostflux, oracle_mod):
    """A jump in the stencil: WENO5 departs from the linear scheme (the
    ramp tests above are not satisfied by any flux)."""
    st = np.array([_state(1.0, [0.1, 0, 0], 1.0)] * 3 + [_state(0.125, [0.1, 0, 0], 0.1)] * 3)
    ref = _face_flux(oracle_mod, list(st), 0, 1, 0, 0)
    F = linear_flux(hostflux, st, 0, 0, 0, 1)
    assert np.abs(F - ref).max() > 1e-3 * np.abs(ref).max()
