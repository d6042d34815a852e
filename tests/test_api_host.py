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


def _rbc_problem(Lz=1.0, bcs=None):
    """The first-order Rayleigh-Benard statement of test/test_subproblem_rk.jl:27-56,
    with stand-in variables (recognition needs names and bases only)."""
    dom = T.ChannelDomain(16, 8, Lx=4.0, Lz=Lz)

    class V:
        def __init__(self, n):
            self.name, self.bases = n, dom.bases
    coords = dom.dist.coords
    ex, ez = T.unit_vector_fields(coords)
    lift_basis = T.derivative_basis(dom.bases[1], 1)
    tau_lift = lambda A: T.lift(A, lift_basis, -1)  # noqa: E731
    u, b, p = V("u"), V("b"), V("p")
    taus = [V(n) for n in ("tau_p", "tau_b1", "tau_b2", "tau_u1", "tau_u2")]
    grad_u = T.grad(u) + ez * tau_lift(taus[3])
    grad_b = T.grad(b) + ez * tau_lift(taus[1])
    pb = T.IVP([p, b, u] + taus)
    T.add_parameters_(pb, kappa=0.01, nu=0.02, Lz=Lz, ez=ez, grad_u=grad_u, grad_b=grad_b, τ_lift=tau_lift)
    T.add_equation_(pb, "trace(grad_u) + tau_p = 0")
    T.add_equation_(pb, "∂t(b) - kappa*div(grad_b) + τ_lift(tau_b2) = -u⋅∇(b)")
    T.add_equation_(pb, "∂t(u) - nu*div(grad_u) + ∇(p) - b*ez + τ_lift(tau_u2) = -u⋅∇(u)")
    for bc in bcs or ("b(z=0) = Lz", "u(z=0) = 0", "b(z=Lz) = 0", "u(z=Lz) = 0", "integ(p) = 0"):
        T.add_bc_(pb, bc)
    return pb


def test_rbc_recognition():
    r = recognize(_rbc_problem(Lz=2.0))
    assert r.family == "rbc2d"
    assert r.extra == {"kappa": 0.01, "nu": 0.02, "b_bottom": 2.0, "b_top": 0.0}
    assert r.names == {"u": "u", "b": "b", "p": "p", "tau_p": "tau_p", "tau_b1": "tau_b1", "tau_b2": "tau_b2",
                       "tau_u1": "tau_u1", "tau_u2": "tau_u2"}
    # free-slip walls, a missing gauge and a wrong buoyancy sign are refused
    with pytest.raises(NotImplementedError):
        recognize(_rbc_problem(bcs=("b(z=0) = Lz", "u(z=0) = 1", "b(z=Lz) = 0", "u(z=Lz) = 0", "integ(p) = 0")))
    with pytest.raises(NotImplementedError):
        recognize(_rbc_problem(bcs=("b(z=0) = Lz", "u(z=0) = 0", "b(z=Lz) = 0", "u(z=Lz) = 0")))
    pb = _rbc_problem()
    pb.equations[2] = pb.equations[2].replace("- b*ez", "+ b*ez")
    with pytest.raises(NotImplementedError):
        recognize(pb)


def test_nusselt_closed_forms():
    """Nu = 1 for conduction; + <uz b>/(kappa dB/Lz) for a manufactured convective state."""
    from tarang_amd import diagnostics as D
    import numpy.polynomial.chebyshev as C
    nx, nz, Lx, Lz, kappa = 16, 33, 4.0, 2.0, 0.01
    x = Lx * np.arange(nx) / nx
    z = Lz * (1 - np.cos(np.pi * np.arange(nz) / (nz - 1))) / 2
    # Clenshaw-Curtis is exact for polynomials of degree nz-1
    w = D.clenshaw_curtis_weights(nz)
    t = -np.cos(np.pi * np.arange(nz) / (nz - 1))
    assert abs(w @ t ** 10 - 2 / 11) < 1e-14 and abs(w.sum() - 2) < 1e-14

    def coeff(g):  # grid (nx, nz) -> (nx/2+1, nz): rfft in x, Chebyshev interpolation in z
        F = np.fft.rfft(g, axis=0)
        V = C.chebvander(t, nz - 1)
        return np.linalg.solve(V, F.T).T
    A, Bamp, kx = 0.3, 0.7, 2 * np.pi / Lx
    b = (Lz - z)[None, :] + A * np.sin(kx * x)[:, None] * np.sin(np.pi * z / Lz)[None, :]
    uz = Bamp * np.sin(kx * x)[:, None] * np.sin(np.pi * z / Lz)[None, :]
    cond = D.nusselt(coeff(np.tile(Lz - z, (nx, 1))), np.zeros((nx // 2 + 1, nz)), nx, Lz, kappa, Lz)
    assert abs(cond - 1.0) < 1e-13
    nu = D.nusselt(coeff(b), coeff(uz), nx, Lz, kappa, Lz)
    assert abs(nu - (1.0 + A * Bamp / 4 / (kappa * Lz / Lz))) < 1e-9


def test_nusselt3_reduces_to_2d():
    """A y-independent 3-D state has the 2-D Nusselt number."""
    from tarang_amd import diagnostics as D
    rng = np.random.default_rng(0)
    nx, ny, nz = 8, 4, 9
    b2 = rng.standard_normal((nx // 2 + 1, nz)) + 1j * rng.standard_normal((nx // 2 + 1, nz))
    w2 = rng.standard_normal((nx // 2 + 1, nz)) + 1j * rng.standard_normal((nx // 2 + 1, nz))
    b2[0] = b2[0].real
    w2[0] = w2[0].real
    b2[nx // 2] = b2[nx // 2].real
    w2[nx // 2] = w2[nx // 2].real
    b3 = np.zeros((nx // 2 + 1, ny, nz), complex)
    w3 = np.zeros_like(b3)
    b3[:, 0] = b2 * ny
    w3[:, 0] = w2 * ny
    a = D.nusselt(b2, w2, nx, 1.0, 0.01, 1.0)
    c = D.nusselt3(b3, w3, nx, ny, 1.0, 0.01, 1.0)
    assert abs(a - c) < 1e-12 * abs(a)


def test_tau_field_signatures():
    """ScalarField(dist, name, bases, dtype) and VectorField(dist, coords, name, bases, dtype)
    (test/test_subproblem_rk.jl:35-40) on a CPU distributor."""
    coords = T.CartesianCoordinates("x", "z")
    dist = T.Distributor(coords, dtype=np.float64)
    xb = T.RealFourier(coords["x"], size=16, bounds=(0.0, 4.0))
    tp = T.ScalarField(dist, "tau_p", (), np.float64)
    tb = T.ScalarField(dist, "tau_b1", (xb,), np.float64)
    tu = T.VectorField(dist, coords, "tau_u1", (xb,), np.float64)
    assert tp.name == "tau_p" and tp.bases == () and tp.dtype == np.float64
    assert tb.bases == (xb,) and tb.name == "tau_b1"
    assert [c.name for c in tu.components] == ["tau_u1_x", "tau_u1_z"] and tu.components[0].bases == (xb,)


def test_netcdf4_checkpoint_refused_with_conversion_hint(tmp_path):
    """A NetCDF-4 (HDF5) file, the default of the reference's nccreate, is refused before any
    field is touched, with the classic-format conversion named."""
    from tarang_amd import checkpoint
    p = tmp_path / "state.nc"
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(ValueError, match="NetCDF-4.*nccopy -k classic"):
        checkpoint.load_field_(None, str(p), "b")
