from .krylov_exp import (
    DEFAULT_MAX_KRYLOV_DIM,
    KrylovExpResult,
    ScaledHamiltonian,
    krylov_exp,
    krylov_exp_impl,
)
from .pchip_torch import PCHIP1D

__all__ = [
    "DEFAULT_MAX_KRYLOV_DIM",
    "KrylovExpResult",
    "PCHIP1D",
    "ScaledHamiltonian",
    "krylov_exp",
    "krylov_exp_impl",
]
