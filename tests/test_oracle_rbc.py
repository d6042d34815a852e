"""Pins of the oracle's 2-D Rayleigh-Benard channel (Chebyshev tau, per-mode systems).

What fixes each part independently of the oracle's own formulas:
  - T-basis derivative: numpy's chebder (Chebyshev series derivative) on [0, Lz]
  - BC rows: T_k(-1) = (-1)^k, T_k(1) = 1; mass matrix = identity on evolved rows
  - assembled L (Laplacians, lifts, pressure, buoyancy, continuity): the textbook
    no-slip onset Ra_c = 1707.762 at k_c = 3.1163 (Chandrasekhar 1961), as the
    root of the largest growth rate of sigma M x = (-L + J) x, J = advection of
    the conduction profile
  - stage loop + padded products: growth of the exact eigenvector at the rate
    exp(sigma T), second order in dt
  - the conduction state is steady; kx = 0 diffusion decays as exp(-kappa pi^2 t)
  - the filtered system sizes (4N+7 at kx = 0, 4N+6 otherwise, the sizes of the reference's
    per-mode subproblems)
"""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest
import scipy.linalg as sl

RA_C, K_C = 1707.762, 3.1163


def make(O, Ra, nx, nz, Lx, Lz=1.0, dt=1e-2, ts="RK222", Pr=1.0):
    kappa, nu = (Ra * Pr) ** -0.5, (Ra / Pr) ** -0.5
    return O.RBC2D(nx, nz, Lx, Lz, kappa, nu, dt, ts)


def zgrid(N, Lz):
    return Lz * (1 - np.cos(np.pi * np.arange(N) / (N - 1))) / 2


def growth_rate(O, Ra, k=K_C, N=24, mode=1):
    s = make(O, Ra, 8, N, 2 * np.pi * mode / k)
    M, L = s.matrices(mode)
    J = np.zeros_like(M)
    J[N + np.arange(N), 3 * N + np.arange(N)] = 1.0   # F_b = -uz d(b0)/dz = +uz for b0 = Lz - z
    w, V = sl.eig(-L + J, M)
    ok = np.isfinite(w)
    i = np.argmax(np.where(ok, w.real, -np.inf))
    return w[i], V[:, i], s


def test_rbc_system_sizes(oracle):
    s = make(oracle, 1e4, 16, 8, 4.0)
    assert s.n(0) == 4 * 8 + 7
    assert all(s.n(m) == 4 * 8 + 6 for m in range(1, 9))


def test_rbc_chebyshev_derivative_block(oracle):
    N, Lz = 12, 1.7
    s = make(oracle, 1e4, 8, N, 3.0, Lz=Lz)
    M, L = s.matrices(2)
    D = L[0:N, 3 * N:4 * N].real   # continuity rows, uz columns: dz
    ref = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        d = C.chebder(e) * (2.0 / Lz)
        ref[:len(d), j] = d
    assert np.abs(D - ref).max() < 1e-12 * np.abs(ref).max()
    # the dx term of the continuity row: i kx on ux
    kx = 2 * np.pi / 3.0 * 2
    assert np.allclose(np.diag(L[0:N, 2 * N:3 * N]), 1j * kx, rtol=0, atol=1e-14)


def test_rbc_bc_and_mass_rows(oracle):
    N = 10
    s = make(oracle, 1e4, 8, N, 4.0)
    M, L = s.matrices(1)
    k = np.arange(N)
    r0 = 4 * N
    assert np.array_equal(L[r0 + 0, N:2 * N].real, (-1.0) ** k)      # b(z=0)
    assert np.array_equal(L[r0 + 3, N:2 * N].real, np.ones(N))       # b(z=Lz)
    assert np.array_equal(L[r0 + 1, 2 * N:3 * N].real, (-1.0) ** k)  # ux(z=0)
    assert np.array_equal(L[r0 + 5, 3 * N:4 * N].real, np.ones(N))   # uz(z=Lz)
    Mr = M.real
    assert np.array_equal(Mr[N:4 * N, N:4 * N], np.eye(3 * N))
    assert np.abs(M).sum() == 3 * N


def test_rbc_critical_rayleigh(oracle):
    """sigma_max changes sign at the textbook onset Ra_c = 1707.762 (k = 3.1163)."""
    lo, _, _ = growth_rate(oracle, 0.99 * RA_C)
    hi, _, _ = growth_rate(oracle, 1.01 * RA_C)
    at, _, _ = growth_rate(oracle, RA_C)
    assert lo.real < -1e-3 and hi.real > 1e-3
    # linear interpolation of the root: within 1e-4 relative of Ra_c
    root = 0.99 * RA_C + (0 - lo.real) * (0.02 * RA_C) / (hi.real - lo.real)
    assert abs(root / RA_C - 1) < 1e-4
    assert abs(at.real) < 1e-5
    assert abs(at.imag) < 1e-10   # exchange of stabilities: real onset


@pytest.mark.parametrize("ts,order", [("RK222", 2), ("RK443", 3), ("SBDF2", 2), ("SBDF1", 1)])
def test_rbc_eigenvector_growth(oracle, ts, order):
    """The stepped solution from the leading eigenvector grows as exp(sigma T) with the
    scheme's order (SBDF2 starts with one SBDF1 step, step_subproblem_multistep.jl)."""
    N, nx, Ra, k = 16, 8, 3000.0, np.pi
    errs = []
    for dt in (4e-2, 2e-2):
        sig, v, _ = growth_rate(oracle, Ra, k=k, N=N)
        s = make(oracle, Ra, nx, N, 2 * np.pi / k, dt=dt, ts=ts)
        z = zgrid(N, 1.0)
        b = oracle.forward_transform(s.d, np.tile(1.0 - z, (nx, 1)))
        amp = 1e-7 / np.abs(v[N:4 * N]).max()
        ux, uz = np.zeros_like(b), np.zeros_like(b)
        b[1] += amp * v[N:2 * N]
        ux[1] = amp * v[2 * N:3 * N]
        uz[1] = amp * v[3 * N:4 * N]
        for var, c in ((oracle.RBC_B, b), (oracle.RBC_UX, ux), (oracle.RBC_UZ, uz)):
            s.set(var, c)
        x0 = np.concatenate([b[1], ux[1], uz[1]])
        T = 2.0
        s.step(int(round(T / dt)))
        x1 = np.concatenate([s.get(oracle.RBC_B)[1], s.get(oracle.RBC_UX)[1], s.get(oracle.RBC_UZ)[1]])
        r = np.vdot(x0, x1) / np.vdot(x0, x0)
        errs.append(abs(r - np.exp(sig.real * T)) / np.exp(sig.real * T))
        # off-eigenvector part: time-discretisation error (and parasitic multistep modes)
        assert np.abs(x1 - r * x0).max() < {1: 5e-3, 2: 1e-4, 3: 1e-5}[order] * np.abs(x1).max()
    assert sig.real > 0.1
    assert errs[0] < {1: 5e-3, 2: 5e-4, 3: 1e-5}[order]
    ratio = errs[0] / errs[1]
    assert 0.75 * 2 ** order < ratio < 1.3 * 2 ** order, ratio


def test_rbc_conduction_is_steady(oracle):
    nx, N, Lz = 16, 12, 1.0
    s = make(oracle, 1e5, nx, N, 4.0)
    z = zgrid(N, Lz)
    b0 = np.tile(Lz - z, (nx, 1))
    s.set(oracle.RBC_B, oracle.forward_transform(s.d, b0))
    s.step(5)
    b = oracle.backward_transform(s.d, s.get(oracle.RBC_B))
    assert np.abs(b - b0).max() < 1e-13
    assert np.abs(s.get(oracle.RBC_UZ)).max() < 1e-13
    assert np.abs(s.get(oracle.RBC_P)).max() == 0.0   # M^+ of the final update


def test_rbc_kx0_diffusion_exact(oracle):
    """A z-only perturbation of b diffuses as exp(-kappa pi^2 t) sin(pi z) (u stays 0)."""
    nx, N, Lz, eps = 8, 32, 1.0, 1e-3
    errs = []
    for dt in (1e-2, 5e-3):
        s = make(oracle, 1e4, nx, N, 4.0, dt=dt)
        kappa = 1e-2
        z = zgrid(N, Lz)
        s.set(oracle.RBC_B, oracle.forward_transform(s.d, np.tile(Lz - z + eps * np.sin(np.pi * z), (nx, 1))))
        T = 1.0
        s.step(int(round(T / dt)))
        b = oracle.backward_transform(s.d, s.get(oracle.RBC_B))
        exact = Lz - z + eps * np.exp(-kappa * np.pi ** 2 * T) * np.sin(np.pi * z)
        errs.append(np.abs(b - exact[None, :]).max() / eps)
    assert errs[0] < 1e-8 and 3.0 < errs[0] / errs[1] < 5.0


def test_rbc_bc_values(oracle):
    """b(z=0) and b(z=Lz) take the Dirichlet values after the implicit stages (kx = 0 mean)."""
    nx, N = 8, 12
    s = make(oracle, 1e4, nx, N, 4.0)
    s.set_bc(2.0, 0.5)
    z = zgrid(N, 1.0)
    s.set(oracle.RBC_B, oracle.forward_transform(s.d, np.tile(2.0 - 1.5 * z, (nx, 1))))
    s.step(3)
    b = oracle.backward_transform(s.d, s.get(oracle.RBC_B))
    assert np.abs(b - (2.0 - 1.5 * z)[None, :]).max() < 1e-13   # the linear profile is steady


def test_rbc_sbdf2_variable_step_order(oracle):
    """Alternating steps h, h/2 (w = 1/2, 2) keep SBDF2 second order: the w-dependent
    coefficients of _sbdf2_coefs (step_subproblem_multistep.jl:544-552)."""
    nx, N, eps, kappa = 8, 24, 1e-3, 1e-2
    errs = []
    for h in (4e-2, 2e-2):
        s = make(oracle, 1e4, nx, N, 4.0, dt=h, ts="SBDF2")
        z = zgrid(N, 1.0)
        s.set(oracle.RBC_B, oracle.forward_transform(s.d, np.tile(1 - z + eps * np.sin(np.pi * z), (nx, 1))))
        t = 0.0
        for i in range(int(round(1.0 / (1.5 * h))) * 2):
            dt = h if i % 2 == 0 else h / 2
            s.set_dt(dt)
            s.step(1)
            t += dt
        assert abs(s.time - t) < 1e-12
        b = oracle.backward_transform(s.d, s.get(oracle.RBC_B))
        exact = 1 - z + eps * np.exp(-kappa * np.pi ** 2 * t) * np.sin(np.pi * z)
        errs.append(np.abs(b - exact[None, :]).max() / eps)
    assert errs[0] < 2e-5 and 3.3 < errs[0] / errs[1] < 4.7, errs


def test_wall_array_bc_values(oracle):
    """Per-mode wall values (orc_rbc_set_bc_values / orc_rbc3_set_bc_values): a constant wall
    given as an array reproduces set_bc exactly, and an SBDF2 state -- one implicit solve per
    step -- takes the wall arrays as its values on z = 0 and z = Lz (pins the mode mapping,
    the unnormalised-rfft scaling and which row is which wall)."""
    import numpy as np
    nx, N, Lx, Lz, dt = 16, 14, 4.0, 1.0, 5e-3
    kappa = nu = 1e-2
    d = oracle.domain([oracle.REALFOURIER, oracle.CHEBYSHEVT], [nx, N], bounds=[(0.0, Lx), (0.0, Lz)])
    z = Lz * (1 - np.cos(np.pi * np.arange(N) / (N - 1))) / 2
    b0 = oracle.forward_transform(d, np.tile(Lz - z, (nx, 1)))
    a = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, "SBDF2")
    c = oracle.RBC2D(nx, N, Lx, Lz, kappa, nu, dt, "SBDF2")
    for s in (a, c):
        s.set(oracle.RBC_B, b0)
    a.set_bc(1.5, -0.25)
    c.set_bc_values(np.fft.rfft(np.full(nx, 1.5)), np.fft.rfft(np.full(nx, -0.25)))
    a.step(3)
    c.step(3)
    assert np.array_equal(a.get(oracle.RBC_B), c.get(oracle.RBC_B))
    x = Lx * np.arange(nx) / nx
    g0, g1 = 1 + 0.3 * np.cos(2 * np.pi * x / Lx), 0.2 * np.sin(4 * np.pi * x / Lx)
    c.set_bc_values(np.fft.rfft(g0), np.fft.rfft(g1))
    c.step(2)
    grid = oracle.backward_transform(d, c.get(oracle.RBC_B))
    assert np.abs(grid[:, 0] - g0).max() < 1e-12 and np.abs(grid[:, -1] - g1).max() < 1e-12
    ny = 8
    d3 = oracle.domain([oracle.REALFOURIER] * 2 + [oracle.CHEBYSHEVT], [nx, ny, N],
                       bounds=[(0.0, Lx), (0.0, 2.0), (0.0, Lz)])
    s3 = oracle.RBC3D(nx, ny, N, Lx, 2.0, Lz, kappa, nu, dt, "SBDF2")
    s3.set(oracle.RBC3_B, oracle.forward_transform(d3, np.tile(Lz - z, (nx, ny, 1))))
    X, Y = np.meshgrid(x, 2.0 * np.arange(ny) / ny, indexing="ij")
    w0, w1 = 1 + 0.2 * np.cos(2 * np.pi * X / Lx) * np.sin(np.pi * Y), 0.1 * np.cos(np.pi * Y)
    s3.set_bc_values(np.fft.fft(np.fft.rfft(w0, axis=0), axis=1), np.fft.fft(np.fft.rfft(w1, axis=0), axis=1))
    s3.step(2)
    g3 = oracle.backward_transform(d3, s3.get(oracle.RBC3_B))
    assert np.abs(g3[:, :, 0] - w0).max() < 1e-12 and np.abs(g3[:, :, -1] - w1).max() < 1e-12
