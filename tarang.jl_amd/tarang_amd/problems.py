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

Unicode spellings of the reference (∂t, Δ, ∇, ⋅) and `dot` are accepted.
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
    problem.parameters.update({k: float(v) for k, v in params.items()})
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
    family: str                    # "diffusion" | "turb2d" | "burgers1d"
    evolved: str                   # evolved variable name
    coeff: float                   # implicit Laplacian coefficient
    names: Dict[str, str] = dfield(default_factory=dict)


def _param(problem: IVP, token: str) -> float:
    try:
        return float(token)
    except ValueError:
        if token not in problem.parameters:
            raise ValueError(f"parameter {token!r} is not defined (use add_parameters_)")
        return problem.parameters[token]


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
    raise NotImplementedError(
        "equation system not supported by the GPU timestep path:\n  " + "\n  ".join(problem.equations) +
        "\nsupported: dt(T) - k*lap(T) = 0; dt(u) - nu*lap(u) = -u*dx(u); the vorticity/streamfunction "
        "system of examples/2d_decaying_turbulence_gpu.jl")
