"""Shared pieces of the state-vector path (mirrors the reference's emu_base)."""

from .constants import DEVICE_COUNT
from .jump_lindblad_operators import compute_noise_from_lindbladians
from .math.brents_root_finding import find_root_brents
from .math.krylov_exp import DEFAULT_MAX_KRYLOV_DIM, krylov_exp
from .math.matmul import matmul_2x2_with_batched
from .pulser_adapter import (
    HamiltonianType,
    InteractionMatrixCallable,
    PulserData,
    SequenceData,
    TIME_TOLERANCE,
    get_target_times,
    interaction_matrix_from_positions,
    is_evaluation_time,
    sequence_data_from_samples,
)
from .utils import (
    apply_measurement_errors,
    get_max_rss_cpu,
    get_max_rss_gpu,
    init_logging,
    unix_like,
)

__all__ = [
    "DEFAULT_MAX_KRYLOV_DIM",
    "DEVICE_COUNT",
    "HamiltonianType",
    "InteractionMatrixCallable",
    "PulserData",
    "SequenceData",
    "TIME_TOLERANCE",
    "apply_measurement_errors",
    "compute_noise_from_lindbladians",
    "find_root_brents",
    "get_max_rss_cpu",
    "get_max_rss_gpu",
    "get_target_times",
    "init_logging",
    "interaction_matrix_from_positions",
    "is_evaluation_time",
    "krylov_exp",
    "matmul_2x2_with_batched",
    "sequence_data_from_samples",
    "unix_like",
]

__version__ = "0.1.0"
