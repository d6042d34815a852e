"""IVP definition and recognition of the supported equation families.

The reference parses arbitrary equation strings into operator trees
(src/core/problems/problem_parsing.jl).  This package accepts the equation
families of the GPU timestep path and refuses anything else loudly:

  diffusion      dt(T) - kappa*lap(T) = 0
  vorticity      dt(q) - nu*lap(q) = -u@grad(q)
                 lap(psi) + tau_psi - q = 0
                 u - skew(grad(psi)) = 0
                 integ(psi) = 0                        (boundary condition)
  burgers        dt(u) - nu*lap(u) = -u*dx(u)
  rbc2d, rbc3d   Rayleigh-Benard on RealFourier(x) [x RealFourier(y)] x ChebyshevT(z), first-order
                 tau form of test/test_subproblem_rk.jl:43-56:
                   trace(grad_u) + tau_p = 0
                   dt(b) - kappa*div(grad_b) + lift(tau_b2) = -u@grad(b)
                   dt(u) - nu*div(grad_u) + grad(p) - b*ez + lift(tau_u2) = -u@grad(u)
                   grad_u = grad(u) + ez*lift(tau_u1), grad_b = grad(b) + ez*lift(tau_b1)
                 with b(z=lo) = B0, b(z=hi) = B1, u(z=lo) = u(z=hi) = 0, integ(p) = 0

Unicode spellings of the reference (∂t, Δ, ∇, ⋅) and `dot` are accepted.
Parameters that are operator expressions (grad_u = grad(u) + ez*τ_lift(tau_u1))
or functions returning one (τ_lift(A) = lift(A, lift_basis, -1)) are substituted
into the equations before recognition.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field as dfield
from typing import Dict, List, Optional


class IVP:
    """Initial value problem over a list of variables (problem_types.jl:30,195)."""

    def __init__(self, variables, namespace: Optional[dict] = None):
        self.variables = list(variables)
        self.parameters: Dict[str, float] = {}
        self.equations: List[str] = []
        self.bcs: List[str] = []

    def var(self, name):
        for v in self.variables:
            if v.name == name:
                return v
        raise KeyError(f"no variable named {name!r} in the problem")


def add_parameters_(problem: IVP, **params) -> IVP:
    """Numbers are stored as floats; operator expressions, unit vectors and
    functions that build expressions are kept for substitution."""
    for k, v in params.items():
        problem.parameters[k] = float(v) if isinstance(v, (int, float)) else v
    return problem


def add_equation_(problem: IVP, eq: str) -> IVP:
    problem.equations.append(eq)
    return problem


def add_bc_(problem: IVP, bc: str) -> IVP:
    problem.bcs.append(bc)
    return problem


_UNICODE = [("∂t", "dt"), ("Δ", "lap"), ("∇⋅", "div"), ("∇", "grad"), ("⋅", "@"), ("·", "@"), ("∫", "integ")]


def canonical(eq: str) -> str:
    s = eq
    for a, b in _UNICODE:
        s = s.replace(a, b)
    s = re.sub(r"\bdot\b", "@", s)
    s = re.sub(r"\s+", "", s)
    s = s.replace("laplacian(", "lap(").replace("integrate(", "integ(")
    return s


_W = r"([A-Za-z_][A-Za-z_0-9]*)"


@dataclass
class Recognized:
    family: str                    # "diffusion" | "turb2d" | "burgers1d" | "rbc2d" | "rbc3d"
    evolved: str                   # evolved variable name
    coeff: float                   # implicit Laplacian coefficient
    names: Dict[str, str] = dfield(default_factory=dict)
    extra: Dict[str, float] = dfield(default_factory=dict)


def _param(problem: IVP, token: str) -> float:
    try:
        return float(token)
    except ValueError:
        if token not in problem.parameters:
            raise ValueError(f"parameter {token!r} is not defined (use add_parameters_)")
        v = problem.parameters[token]
        if not isinstance(v, float):
            raise ValueError(f"parameter {token!r} is not a number")
        return v


def expand(problem: IVP, eq: str) -> str:
    """Substitute expression-valued and function-valued parameters into `eq`
    (canonical text in, canonical text out)."""
    from .operators import Expr
    s = canonical(eq)
    for _ in range(4):  # nested substitutions (a parameter expression naming another)
        before = s
        for name, v in problem.parameters.items():
            key = canonical(name)
            if callable(v) and not isinstance(v, Expr):
                def repl(m, fn=v):
                    return canonical(str(fn(Expr(m.group(1)))))
                s = re.sub(rf"(?<![A-Za-z_0-9]){re.escape(key)}\(([^()]*)\)", repl, s)
            elif isinstance(v, Expr):
                s = re.sub(rf"(?<![A-Za-z_0-9]){re.escape(key)}(?![A-Za-z_0-9(])", lambda m, t=v.text: t, s)
        if s == before:
            break
    return s


def _bc_value(problem: IVP, token: str):
    """A BC right-hand side: a number, or a parameter holding a number or a grid array
    of wall values (a space-dependent BC, the reference's ArrayOperator,
    src/core/subsystems/subproblem_rhs.jl:30-58)."""
    v = problem.parameters.get(token)
    if v is not None and not isinstance(v, float) and hasattr(v, "shape"):
        return v
    return _param(problem, token)


def recognize(problem: IVP) -> Recognized:
    eqs = [canonical(e) for e in problem.equations]
    bcs = [canonical(b) for b in problem.bcs]
    # single-equation families
    if len(eqs) == 1:
        m = re.fullmatch(rf"dt\({_W}\)-{_W}\*lap\(\1\)=0", eqs[0])
        if m:
            return Recognized("diffusion", m.group(1), _param(problem, m.group(2)))
        m = re.fullmatch(rf"dt\({_W}\)-{_W}\*lap\(\1\)=-\1\*(?:dx\(\1\)|d\(\1,x\))", eqs[0])
        if m:
            return Recognized("burgers1d", m.group(1), _param(problem, m.group(2)))
    # vorticity / streamfunction system
    if len(eqs) == 3:
        mq = re.fullmatch(rf"dt\({_W}\)-{_W}\*lap\(\1\)=-{_W}@grad\(\1\)", eqs[0])
        mp = re.fullmatch(rf"lap\({_W}\)\+{_W}-{_W}=0", eqs[1])
        mu = re.fullmatch(rf"{_W}-skew\(grad\({_W}\)\)=0", eqs[2])
        if mq and mp and mu:
            q, nu, u = mq.group(1), mq.group(2), mq.group(3)
            psi, tau, q2 = mp.group(1), mp.group(2), mp.group(3)
            u2, psi2 = mu.group(1), mu.group(2)
            gauge = any(re.fullmatch(rf"integ\({re.escape(psi)}\)=0", b) for b in bcs)
            if q2 == q and u2 == u and psi2 == psi and gauge:
                return Recognized("turb2d", q, _param(problem, nu), {"psi": psi, "u": u, "tau": tau})
    rec = _recognize_rbc2d(problem)
    if rec is not None:
        return rec
    raise NotImplementedError(
        "equation system not supported by the GPU timestep path:\n  " + "\n  ".join(problem.equations) +
        "\nsupported: dt(T) - k*lap(T) = 0; dt(u) - nu*lap(u) = -u*dx(u); the vorticity/streamfunction "
        "system of examples/2d_decaying_turbulence_gpu.jl; the first-order tau form of 2-D "
        "Rayleigh-Benard of test/test_subproblem_rk.jl")


def _recognize_rbc2d(problem: IVP) -> Optional[Recognized]:
    if len(problem.equations) != 3:
        return None
    eqs = [expand(problem, e) for e in problem.equations]
    L = r"lift\({w},-1\)"
    e1 = re.fullmatch(rf"trace\(grad\({_W}\)\+ez\*{L.format(w=_W)}\)\+{_W}=0", eqs[0])
    e2 = re.fullmatch(rf"dt\({_W}\)-{_W}\*div\(grad\(\1\)\+ez\*{L.format(w=_W)}\)\+{L.format(w=_W)}"
                      rf"=-{_W}@grad\(\1\)", eqs[1])
    e3 = re.fullmatch(rf"dt\({_W}\)-{_W}\*div\(grad\(\1\)\+ez\*{L.format(w=_W)}\)\+grad\({_W}\)-{_W}\*ez"
                      rf"\+{L.format(w=_W)}=-\1@grad\(\1\)", eqs[2])
    if not (e1 and e2 and e3):
        return None
    u, tau_u1, tau_p = e1.groups()
    b, kappa, tau_b1, tau_b2, u2 = e2.groups()
    u3, nu, tau_u1b, p, b3, tau_u2 = e3.groups()
    if not (u2 == u3 == u and b3 == b and tau_u1b == tau_u1):
        return None
    names = {"u": u, "b": b, "p": p, "tau_p": tau_p, "tau_b1": tau_b1, "tau_b2": tau_b2,
             "tau_u1": tau_u1, "tau_u2": tau_u2}
    bcs = {}
    for bc in problem.bcs:
        c = canonical(bc)
        m = re.fullmatch(rf"{_W}\(z=([^)]+)\)=(.+)", c)
        if m:
            bcs[(m.group(1), _bc_value(problem, m.group(2)))] = _bc_value(problem, m.group(3))
            continue
        m = re.fullmatch(rf"integ\({_W}\)=0", c)
        if m and m.group(1) == p:
            bcs["integ"] = True
            continue
        raise NotImplementedError(f"boundary condition {bc!r} is not supported by the channel GPU path")
    zb = problem.var(b).bases[-1]
    lo, hi = zb.bounds
    need = [(b, lo), (b, hi), (u, lo), (u, hi)]
    for key in need:
        if key not in bcs:
            raise NotImplementedError(f"missing boundary condition {key[0]}(z={key[1]:g})")
    if any(hasattr(bcs[k], "shape") for k in ((u, lo), (u, hi))) or bcs[(u, lo)] != 0.0 or bcs[(u, hi)] != 0.0:
        raise NotImplementedError("the channel GPU path supports no-slip walls u = 0 only")
    if "integ" not in bcs:
        raise NotImplementedError(f"missing gauge condition integ({p}) = 0")
    ndim = len(problem.var(b).bases)
    rec = Recognized("rbc3d" if ndim == 3 else "rbc2d", b, _param(problem, kappa), names)
    rec.extra = {"kappa": _param(problem, kappa), "nu": _param(problem, nu),
                 "b_bottom": bcs[(b, lo)], "b_top": bcs[(b, hi)]}
    return rec
