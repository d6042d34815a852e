"""Host-side logic of the API mirror (no GPU needed)."""
import math

import numpy as np
import pytest

import tarang_amd as T
from tarang_amd.problems import canonical, recognize


def test_domain_shapes_and_grids():
    d = T.PeriodicDomain(16, 8, L=(2.0, 4.0))
    assert d.grid_shape == (16, 8)
    assert d.coeff_shape == (9, 8)
    assert np.allclose(d.bases[0].grid(), np.arange(16) * 2.0 / 16)
    kf = d.bases[1].wavenumbers_fft() / (2 * math.pi / 4.0)
    assert list(kf) == [0, 1, 2, 3, -4, -3, -2, -1]
    assert list(d.bases[0].wavenumbers_rfft() / math.pi) == list(range(9))
    c = T.ChannelDomain(32, 9, Lx=4.0, Lz=1.0)
    z = c.bases[1].grid()
    assert abs(z[0]) < 1e-15 and abs(z[-1] - 1.0) < 1e-15 and np.all(np.diff(z) > 0)
    c3 = T.ChannelDomain3D(8, 8, 9)
    assert c3.bases[2].bounds == (-1.0, 1.0)  # reference quirk: Lz ignored


def test_cpu_arch_refuses():
    d = T.PeriodicDomain(16)
    with pytest.raises(T.GPUTransformUnsupported):
        d.device()


def test_equation_recognition():
    class V:
        def __init__(self, n):
            self.name = n
    p = T.IVP([V("q"), V("psi"), V("u"), V("tau_psi")])
    T.add_parameters_(p, nu=1e-4)
    T.add_equation_(p, "∂t(q) - nu*Δ(q) = -u⋅∇(q)")
    T.add_equation_(p, "Δ(psi) + tau_psi - q = 0")
    T.add_equation_(p, "u - skew(grad(psi)) = 0")
    T.add_bc_(p, "integ(psi) = 0")
    r = recognize(p)
    assert r.family == "turb2d" and r.evolved == "q" and r.coeff == 1e-4 and r.names["psi"] == "psi"

    p = T.IVP([V("T")])
    T.add_parameters_(p, kappa=0.01)
    T.add_equation_(p, "dt(T) - kappa*lap(T) = 0")
    assert recognize(p).family == "diffusion"

    p = T.IVP([V("u")])
    T.add_parameters_(p, nu=0.01)
    T.add_equation_(p, "dt(u) - nu*lap(u) = -u*dx(u)")
    assert recognize(p).family == "burgers1d"

    p = T.IVP([V("u")])
    T.add_equation_(p, "dt(u) - nu*lap(u) = -u*dx(u)")
    with pytest.raises(ValueError):  # undefined parameter
        recognize(p)

    p = T.IVP([V("u")])
    T.add_equation_(p, "dt(u) + dx(u)*dx(u) = 0")
    with pytest.raises(NotImplementedError):
        recognize(p)


def test_canonical_forms():
    assert canonical("u dot grad(b)") == "u@grad(b)"
    assert canonical(" ∂t(q) - ν*Δ(q) = 0") == "dt(q)-ν*lap(q)=0"


def test_timestepper_codes_match_header():
    assert (T.RK222.code, T.RK443.code, T.SBDF2.code, T.RK111.code, T.SBDF1.code) == (0, 1, 2, 3, 4)
    assert T.DiagonalIMEX_SBDF2.code == T.SBDF2.code
