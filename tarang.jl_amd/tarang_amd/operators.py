"""Symbolic operator helpers used to state the supported equation families.

The reference builds operator trees (src/core/operators/*) and substitutes them
into equation strings through problem parameters (`add_parameters!(problem,
grad_u=grad_u, τ_lift=τ_lift, ...)`, test/test_subproblem_rk.jl:43-49).  Here an
expression is kept as its canonical text; recognition of the equation family
(problems.recognize) works on the text after substitution.  Only the operators
those families use are provided: grad, lift, unit vectors, derivative_basis.
"""
from __future__ import annotations


def _text(x) -> str:
    if isinstance(x, Expr):
        return x.text
    if hasattr(x, "name") and isinstance(getattr(x, "name"), str):
        return x.name
    if isinstance(x, (int, float)):
        return repr(x)
    return str(x)


class Expr:
    """An operator expression held as canonical text (no spaces)."""

    def __init__(self, text: str):
        self.text = text

    def __add__(self, o):
        return Expr(f"{self.text}+{_text(o)}")

    def __radd__(self, o):
        return Expr(f"{_text(o)}+{self.text}")

    def __sub__(self, o):
        return Expr(f"{self.text}-{_text(o)}")

    def __rsub__(self, o):
        return Expr(f"{_text(o)}-{self.text}")

    def __mul__(self, o):
        return Expr(f"{self.text}*{_text(o)}")

    def __rmul__(self, o):
        return Expr(f"{_text(o)}*{self.text}")

    def __neg__(self):
        return Expr(f"-{self.text}")

    def __repr__(self):
        return f"Expr({self.text!r})"

    def __str__(self):
        return self.text


def grad(f) -> Expr:
    """Gradient operator (src/core/operators/*; `∇`)."""
    return Expr(f"grad({_text(f)})")


class DerivativeBasis:
    """derivative_basis(basis, order): the basis a lift is stated in.  The per-mode
    lift column is the unit vector at Chebyshev index N-1 whatever this basis is
    (src/core/operators/matrices/matrices_subproblem_operators.jl:218-261)."""

    def __init__(self, basis, order: int):
        self.basis = basis
        self.order = int(order)


def derivative_basis(basis, order: int = 1) -> DerivativeBasis:
    return DerivativeBasis(basis, order)


def lift(tau, basis, n: int) -> Expr:
    """lift(tau, basis, n): tau times the n-th mode (n = -1: the last) of `basis`."""
    return Expr(f"lift({_text(tau)},{int(n)})")


def unit_vector_fields(coords, dist=None):
    """(e_x, e_y, ...) unit vectors named e<coordinate>."""
    return tuple(Expr(f"e{n}") for n in coords.names)
