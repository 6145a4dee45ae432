"""Host-side mirrors of the reference's ``Mesh``, ``Point`` and ``Solid``
(src/boot/needs/types.jl:38-167).

Arrays are numpy arrays in Fortran (column-major) order with the reference's
shapes, so ``mpts.x[i, p]`` is Julia's ``mpts.x[i+1, p+1]`` and the raw memory
is exactly what a Julia ``ccall`` would pass to the C ABI.  Fields that the
reference keeps but never reads on the step path (``Mᵢⱼ``, ``mesh.D``,
``mesh.f``, ``mesh.bij``, ``mesh.Δu``, ``s.u``, ``s.p``, ``s.ϵpV``,
``s.∇uᵢⱼ``, APIC ``Bᵢⱼ``/``Dᵢⱼ``, non-local ``e2p``/``p2p``) are not allocated.
The materialised stencil (``ϕ∂ϕ``, ``Δnp``, ``p2n``, ``p2e``) is only
allocated on request, for the CPU oracle: the device never stores it.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def fzeros(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


@dataclass
class Boundary:
    status: np.ndarray                 # Bool [dim, nno]


@dataclass
class Mesh:
    dim: int
    nel: np.ndarray                    # [dim+1] (per axis + total)
    nno: np.ndarray                    # [dim+1]
    nn: int
    L: np.ndarray
    h: np.ndarray
    x0: np.ndarray                     # x₀
    x: np.ndarray                      # [dim, nno]
    m: np.ndarray                      # mᵢ [nno]
    oobf: np.ndarray                   # [dim, nno]
    a: np.ndarray
    p: np.ndarray
    v: np.ndarray
    dJ: np.ndarray                     # ΔJ [dim, nno]
    e2n: np.ndarray                    # [nn, nel] int64, 1-based, 0 = absent
    xB: np.ndarray                     # [dim, 2]
    bcs: Boundary


@dataclass
class Solid:
    v: np.ndarray                      # [dim, nmp]
    rho0: np.ndarray                   # ρ₀
    rho: np.ndarray                    # ρ
    c0: np.ndarray                     # c₀
    cr: np.ndarray                     # cᵣ
    phi: np.ndarray                    # ϕ
    dlam: np.ndarray                   # Δλ
    epII: np.ndarray                   # ϵpII [2, nmp]
    sig: np.ndarray                    # σᵢ [nstr, nmp]
    tau: np.ndarray                    # τᵢ [nstr, nmp]
    gradv: np.ndarray                  # ∇vᵢⱼ [dim, dim, nmp]
    dF: np.ndarray                     # ΔFᵢⱼ
    F: np.ndarray                      # Fᵢⱼ
    b: np.ndarray                      # bᵢⱼ
    eps: np.ndarray                    # ϵᵢⱼ
    omg: np.ndarray                    # ωᵢⱼ


@dataclass
class Point:
    ndim: int
    nmp: int
    vmax: np.ndarray                   # [dim]
    x: np.ndarray                      # [dim, nmp]
    l0: np.ndarray                     # ℓ₀
    l: np.ndarray                      # ℓ
    n0: np.ndarray
    n: np.ndarray
    Om0: np.ndarray                    # Ω₀
    Om: np.ndarray                     # Ω
    dJ: np.ndarray                     # ΔJ
    J: np.ndarray
    s: Solid
    z0: np.ndarray = field(default=None)
    Bij: np.ndarray | None = None      # Bᵢⱼ [dim, dim, nmp], APIC state (types.jl:81-83)
    # materialised stencil, CPU oracle only
    phi: np.ndarray | None = None      # ϕ∂ϕ [nn, nmp, dim+1]
    dnp: np.ndarray | None = None      # Δnp [nn, dim, nmp]
    Dij: np.ndarray | None = None      # Dᵢⱼ [dim, dim, nmp] (APIC)
    p2e: np.ndarray | None = None
    p2n: np.ndarray | None = None

    def ensure_stencil(self, nn: int):
        """Allocate ϕ∂ϕ / Δnp / p2n / p2e (setup_mpts.jl:86-97)."""
        if self.phi is None:
            d = self.ndim
            self.phi = fzeros((nn, self.nmp, d + 1))
            self.dnp = fzeros((nn, d, self.nmp))
            self.p2e = np.zeros(self.nmp, dtype=np.int64)
            self.p2n = np.zeros((nn, self.nmp), dtype=np.int64, order="F")
        if self.Dij is None:
            self.Dij = fzeros((self.ndim, self.ndim, self.nmp))
        if self.Bij is None:
            self.Bij = fzeros((self.ndim, self.ndim, self.nmp))
        return self
