"""Host-side mirror of Tarang.jl's public API for the GPU timestep path.

Names follow the reference; Julia's mutating `!` suffix becomes a trailing
underscore (`forward_transform!` -> `forward_transform_`).  Device memory is
torch tensors on `cuda:<n>`; all arithmetic on the hot path runs in
libtarang_hip.so.  Arrays are exposed in the reference's logical index order
(`field["g"][ix, iy]`), backed by column-major storage exactly like Julia's.

Reference surface mirrored here:
  architectures   src/core/architectures.jl:60-131
  bases/domains   src/core/basis/*, src/core/domain.jl:101-145,
                  src/extras/quick_domains.jl:33-143
  fields          src/core/field/field_types.jl:96-180, quick_domains.jl:192-307
  transforms      src/core/transforms/transform_gpu.jl:220-258
  problems        src/core/problems/problem_types.jl:30,195,506,748
  solver          src/core/solvers/solver_types.jl:220-553, solver_stepping.jl:24-206
"""
from __future__ import annotations

import ctypes
import math
import re
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


class GPUTransformUnsupported(RuntimeError):
    """The requested operation has no GPU implementation (never falls back to the CPU)."""


# --------------------------------------------------------------------------
# architectures
# --------------------------------------------------------------------------
class CPU:
    """The reference's CPU architecture.  This package is the GPU backend only."""

    def __repr__(self):
        return "CPU()"


class GPU:
    """GPU architecture (one MI355X device)."""

    def __init__(self, device: int = 0):
        self.device = int(device)

    def __repr__(self):
        return f"GPU({self.device})"


def _is_gpu(arch) -> bool:
    return isinstance(arch, GPU)


class _Device:
    """One library context per device; the context runs on its own HIP stream and
    is synchronised with torch at every hand-over."""

    _ctxs: Dict[int, "_Device"] = {}

    def __init__(self, index: int):
        if torch is None or not torch.cuda.is_available():
            raise GPUTransformUnsupported("GPU(): no HIP device is visible to this process")
        self.index = index
        self.torch_device = torch.device("cuda", index)
        h = ctypes.c_void_p()
        _lib.call("tarang_ctx_create", index, ctypes.byref(h))
        self.handle = h

    @classmethod
    def get(cls, index: int) -> "_Device":
        if index not in cls._ctxs:
            cls._ctxs[index] = _Device(index)
        return cls._ctxs[index]

    def before(self):
        """Make torch's pending work on this device visible to the library."""
        torch.cuda.synchronize(self.torch_device)

    def after(self):
        """Make the library's work visible to torch."""
        _lib.call("tarang_ctx_sync", self.handle)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


# --------------------------------------------------------------------------
# coordinates, bases, distributor, domains
# --------------------------------------------------------------------------
class Coordinate:
    def __init__(self, name: str):
        self.name = name

    def __repr__(self):
        return f"Coordinate({self.name!r})"


class CartesianCoordinates:
    def __init__(self, *names: str):
        self.names = tuple(names)
        self.coords = {n: Coordinate(n) for n in names}

    def __getitem__(self, name: str) -> Coordinate:
        return self.coords[name]

    def __len__(self):
        return len(self.names)


def _np_dtype(dtype):
    if dtype in (np.float32, "float32", "Float32") or (torch is not None and dtype == torch.float32):
        return np.float32
    if dtype in (np.float64, float, "float64", "Float64", None) or (torch is not None and dtype == torch.float64):
        return np.float64
    raise ValueError(f"unsupported dtype {dtype!r} (Float32 or Float64)")


class Distributor:
    def __init__(self, coords: CartesianCoordinates, dtype=np.float64, architecture=None, device=None, mesh=None):
        if mesh is not None and tuple(mesh) not in ((), (1,)):
            raise GPUTransformUnsupported("distributed meshes are not supported: one GPU, no collectives")
        self.coords = coords
        self.dtype = _np_dtype(dtype)
        self.architecture = device if device is not None else (architecture if architecture is not None else CPU())


class _Basis:
    kind = None

    def __init__(self, coord: Coordinate, size: int, bounds=(0.0, 2 * math.pi), dealias: float = 1.0, dtype=None):
        if int(size) < 1:
            raise ValueError("basis size must be positive")
        self.coord = coord
        self.size = int(size)
        self.bounds = (float(bounds[0]), float(bounds[1]))
        self.dealias = float(dealias)

    @property
    def length(self):
        return self.bounds[1] - self.bounds[0]


class RealFourier(_Basis):
    """RealFourier basis: grid x_j = a + j L/N (src/core/basis/basis_interface.jl:81-92)."""

    kind = _lib.REALFOURIER

    def grid(self) -> np.ndarray:
        a = self.bounds[0]
        dx = self.length / self.size
        return np.array([a + i * dx for i in range(self.size)])

    def wavenumbers_rfft(self) -> np.ndarray:
        """[0, 1, ..., N/2] * 2pi/L (src/core/basis/basis_wavenumbers.jl:69-79)."""
        return (2 * math.pi / self.length) * np.arange(self.size // 2 + 1, dtype=np.float64)

    def wavenumbers_fft(self) -> np.ndarray:
        """FFT order [0..N/2-1, -N/2..-1] * 2pi/L (basis_wavenumbers.jl:94-108)."""
        N = self.size
        k = [i if i < (N + 1) // 2 else i - N for i in range(N)]
        return (2 * math.pi / self.length) * np.array(k, dtype=np.float64)


class ChebyshevT(_Basis):
    """ChebyshevT basis on Gauss-Lobatto points z_j = -cos(pi j/(N-1)) mapped to the
    bounds (src/core/basis/basis_interface.jl:107-121)."""

    kind = _lib.CHEBYSHEVT

    def grid(self) -> np.ndarray:
        N = self.size
        if N == 1:
            x = np.array([0.0])
        else:
            x = np.array([-math.cos(math.pi * k / (N - 1)) for k in range(N)])
        a, b = self.bounds
        return (b - a) / 2 * x + (b + a) / 2


class Domain:
    """Tensor-product domain; builds the transform plan on construction
    (src/core/domain.jl:101-145)."""

    def __init__(self, dist: Distributor, bases: Sequence[_Basis]):
        self.dist = dist
        self.bases = tuple(bases)
        self.ndim = len(self.bases)
        if not 1 <= self.ndim <= 3:
            raise ValueError("domains are 1-D, 2-D or 3-D")
        self._plan = None

    @property
    def dtype(self):
        return self.dist.dtype

    @property
    def arch(self):
        return self.dist.architecture

    @property
    def grid_shape(self):
        return tuple(b.size for b in self.bases)

    @property
    def coeff_shape(self):
        """(n0/2+1, n1, ...) complex: rfft on the first RealFourier axis
        (src/core/transforms/transform_layout.jl:41-100, src/core/domain.jl:485-497)."""
        s = [b.size for b in self.bases]
        if isinstance(self.bases[0], RealFourier):
            s[0] = s[0] // 2 + 1
        return tuple(s)

    def device(self) -> _Device:
        if not _is_gpu(self.arch):
            raise GPUTransformUnsupported(
                "this package is the GPU backend: build the domain with arch=GPU() "
                "(the CPU path is the reference's own)")
        return _Device.get(self.arch.device)

    def plan(self):
        if self._plan is None:
            dev = self.device()
            d = _lib.DomainDesc()
            d.ndim = self.ndim
            for a, b in enumerate(self.bases):
                d.kind[a] = b.kind
                d.n[a] = b.size
                d.lo[a], d.hi[a] = b.bounds
                d.dealias[a] = b.dealias
            d.dtype = _lib.F32 if self.dtype == np.float32 else _lib.F64
            h = ctypes.c_void_p()
            _lib.call("tarang_plan_create", dev.handle, ctypes.byref(d), ctypes.byref(h))
            self._plan = h
        return self._plan

    def grid(self, axis: int) -> np.ndarray:
        return self.bases[axis].grid()


def get_grid_coordinates(domain: Domain, on_device: bool = False) -> Dict[str, object]:
    out = {}
    for b in domain.bases:
        g = b.grid()
        out[b.coord.name] = torch.as_tensor(g, device=domain.device().torch_device) if on_device else g
    return out


def PeriodicDomain(*N: int, L=None, dtype=np.float64, device=None, arch=None, dealias: float = 1.5) -> Domain:
    """src/extras/quick_domains.jl:33-58"""
    arch = device if device is not None else (arch if arch is not None else CPU())
    nd = len(N)
    if not 1 <= nd <= 3:
        raise ValueError(f"PeriodicDomain supports 1D, 2D, or 3D (got {nd}D)")
    Ls = tuple(L) if L is not None else (2 * math.pi,) * nd
    if len(Ls) != nd:
        raise ValueError("Length of L must match number of dimensions")
    names = ("x", "y", "z")[:nd]
    coords = CartesianCoordinates(*names)
    dist = Distributor(coords, dtype=dtype, architecture=arch)
    bases = [RealFourier(coords[names[i]], size=int(N[i]), bounds=(0.0, float(Ls[i])), dealias=dealias)
             for i in range(nd)]
    return Domain(dist, bases)


def ChannelDomain(Nx: int, Nz: int, Lx: float = 2 * math.pi, Lz: float = 2.0, dtype=np.float64, arch=None,
                  device=None, dealias: float = 1.5) -> Domain:
    """src/extras/quick_domains.jl:109-121: RealFourier x on [0, Lx], ChebyshevT z on [0, Lz]."""
    arch = device if device is not None else (arch if arch is not None else CPU())
    coords = CartesianCoordinates("x", "z")
    dist = Distributor(coords, dtype=dtype, architecture=arch)
    xb = RealFourier(coords["x"], size=Nx, bounds=(0.0, float(Lx)), dealias=dealias)
    zb = ChebyshevT(coords["z"], size=Nz, bounds=(0.0, float(Lz)), dealias=dealias)
    return Domain(dist, (xb, zb))


def ChannelDomain3D(Nx: int, Ny: int, Nz: int, Lx: float = 4 * math.pi, Ly: float = 2 * math.pi, Lz: float = 2.0,
                    dtype=np.float64, arch=None, device=None, dealias: float = 1.5) -> Domain:
    """src/extras/quick_domains.jl:128-143.  As in the reference, z spans (-1, 1)
    whatever Lz is; build Domain(dist, (xb, yb, zb)) with explicit bounds otherwise."""
    arch = device if device is not None else (arch if arch is not None else CPU())
    coords = CartesianCoordinates("x", "y", "z")
    dist = Distributor(coords, dtype=dtype, architecture=arch)
    xb = RealFourier(coords["x"], size=Nx, bounds=(0.0, float(Lx)), dealias=dealias)
    yb = RealFourier(coords["y"], size=Ny, bounds=(0.0, float(Ly)), dealias=dealias)
    zb = ChebyshevT(coords["z"], size=Nz, bounds=(-1.0, 1.0), dealias=dealias)
    return Domain(dist, (xb, yb, zb))


# --------------------------------------------------------------------------
# fields
# --------------------------------------------------------------------------
def _torch_real(np_dtype):
    return torch.float32 if np_dtype == np.float32 else torch.float64


def _torch_cplx(np_dtype):
    return torch.complex64 if np_dtype == np.float32 else torch.complex128


class ScalarField:
    """Scalar field with a grid buffer, a coefficient buffer and a layout flag
    (`current_layout` in {'g', 'c'}; src/core/field/field_types.jl:96)."""

    def __init__(self, domain, name: str = "", dtype=None, bases=None):
        if isinstance(domain, Distributor):  # ScalarField(dist, name, bases, dtype): constant/tau field
            if isinstance(dtype, (tuple, list)):  # positional (bases, dtype) after the name
                bases, dtype = dtype, bases
            if not isinstance(name, str):  # ScalarField(dist, bases, name) spelling
                name, bases = bases, name
            self.domain = None
            self.dist = domain
            self.name = name
            self.bases = tuple(bases or ())
            self.dtype = _np_dtype(dtype if dtype is not None else domain.dtype)
            self.current_layout = "c"
            self._g = self._c = None
            # coefficient storage of a tau variable: one value per x mode (or one value)
            if _is_gpu(domain.architecture):
                shape = [self.bases[0].size // 2 + 1] + [b.size for b in self.bases[1:]] if self.bases else [1]
                dev = _Device.get(domain.architecture.device)
                self._c = torch.zeros(tuple(reversed(shape)), dtype=_torch_cplx(self.dtype), device=dev.torch_device)
            return
        self.domain = domain
        self.dist = domain.dist
        self.name = name
        self.bases = domain.bases
        self.dtype = _np_dtype(dtype if dtype is not None else domain.dtype)
        if self.dtype != domain.dtype:
            raise ValueError("field dtype must match the domain (Distributor) dtype")
        dev = domain.device()
        rshape = tuple(reversed(domain.grid_shape))
        cshape = tuple(reversed(domain.coeff_shape))
        self._g = torch.zeros(rshape, dtype=_torch_real(self.dtype), device=dev.torch_device)
        self._c = torch.zeros(cshape, dtype=_torch_cplx(self.dtype), device=dev.torch_device)
        self.current_layout = "g"

    # logical (reference-order) views of the column-major storage
    @staticmethod
    def _logical(t):
        return t.permute(*reversed(range(t.dim())))

    @property
    def grid_data(self):
        return self._logical(self._g)

    @property
    def coeff_data(self):
        return self._logical(self._c)

    def __getitem__(self, layout: str):
        ensure_layout_(self, layout)
        return self.grid_data if layout == "g" else self.coeff_data

    def __setitem__(self, layout: str, value):
        if layout not in ("g", "c"):
            raise KeyError(layout)
        dst = self.grid_data if layout == "g" else self.coeff_data
        src = torch.as_tensor(np.asarray(value) if not torch.is_tensor(value) else value)
        dst.copy_(src.to(device=dst.device, dtype=dst.dtype).reshape(dst.shape))
        self.current_layout = layout

    def __repr__(self):
        return f"ScalarField({self.name!r}, layout={self.current_layout})"


class VectorField:
    """Vector field: one ScalarField per coordinate (src/core/field/field_types.jl:152)."""

    def __init__(self, domain, name="", dtype=None, bases=None, dtype_=None):
        if isinstance(domain, Distributor):  # VectorField(dist, coords, name, bases, dtype): tau vector
            dist, coords = domain, name
            vname, vbases, vdtype = dtype, bases, dtype_
            self.domain = None
            self.name = vname
            self.components = [ScalarField(dist, f"{vname}_{n}", vbases, vdtype) for n in coords.names]
            return
        self.domain = domain
        self.name = name
        names = domain.dist.coords.names
        self.components = [ScalarField(domain, f"{name}_{n}", dtype=dtype) for n in names]

    def __getitem__(self, i):
        return self.components[i]

    def __len__(self):
        return len(self.components)


def forward_transform_(field: ScalarField) -> ScalarField:
    """grid -> coefficients on the device (transform_gpu.jl:220-258)."""
    dev = field.domain.device()
    dev.before()
    _lib.call("tarang_forward_transform", field.domain.plan(), _ptr(field._g), _ptr(field._c))
    dev.after()
    field.current_layout = "c"
    return field


def backward_transform_(field: ScalarField) -> ScalarField:
    """coefficients -> grid on the device."""
    dev = field.domain.device()
    dev.before()
    _lib.call("tarang_backward_transform", field.domain.plan(), _ptr(field._c), _ptr(field._g))
    dev.after()
    field.current_layout = "g"
    return field


def ensure_layout_(field: ScalarField, layout: str) -> ScalarField:
    if layout not in ("g", "c"):
        raise ValueError(f"layout must be 'g' or 'c', got {layout!r}")
    if field._g is None or field.current_layout == layout:
        return field
    return forward_transform_(field) if layout == "c" else backward_transform_(field)


def padded_multiply(a: ScalarField, b: ScalarField, out: Optional[ScalarField] = None) -> ScalarField:
    """3/2-dealiased product a*b on the grid (evaluate_padded_multiply,
    src/core/nonlinear/nonlinear_padding.jl:359-437)."""
    if a.domain is not b.domain:
        raise ValueError("Cannot multiply fields with different bases")
    out = out or ScalarField(a.domain, f"{a.name}*{b.name}")
    ensure_layout_(a, "g")
    ensure_layout_(b, "g")
    dev = a.domain.device()
    dev.before()
    _lib.call("tarang_padded_multiply", a.domain.plan(), _ptr(a._g), _ptr(b._g), _ptr(out._g))
    dev.after()
    out.current_layout = "g"
    return out


def set_(field: ScalarField, f) -> ScalarField:
    """set!(field, f): f(x[, y[, z]]) evaluated on the grid (quick_domains.jl:264-307)."""
    coords = np.meshgrid(*[b.grid() for b in field.bases], indexing="ij")
    field["g"] = f(*coords) if callable(f) else f
    return field
