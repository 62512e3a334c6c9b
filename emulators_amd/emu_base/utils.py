"""Logging, memory and readout-error helpers.

Behaviour follows the reference's emu_base/utils.py:14-135: the
"emulators" logger, the peak-memory probes, and the readout-error
resampling of a bag of bitstrings (one ``random.random()`` draw per
character, bitstrings in the bag's order, so a seeded run draws the same
errors as the reference's)."""

from __future__ import annotations

import logging
import os
import random
import sys
from collections import Counter
from pathlib import Path

import torch

unix_like = os.name != "nt"  # platforms with the resource module (public name of emu_base)
LOGGER_NAME = "emulators"
_MESSAGE_ONLY = logging.Formatter("%(message)s")


def _handler_for(log_file: Path | None) -> logging.Handler:
    # stdout by default; a log file is truncated on every (re)configuration
    if log_file is None:
        return logging.StreamHandler(sys.stdout)
    return logging.FileHandler(log_file, mode="w")


def init_logging(log_level: int, log_file: Path | None) -> logging.Logger:
    """(Re)configure the "emulators" logger: it passes everything to a single
    handler of level ``log_level`` and never forwards to the root logger."""
    logger = logging.getLogger(LOGGER_NAME)
    logger.propagate = False
    logger.setLevel(logging.DEBUG)  # the handler's level does the filtering
    handler = _handler_for(log_file)
    handler.setLevel(log_level)
    handler.setFormatter(_MESSAGE_ONLY)
    while logger.handlers:
        logger.removeHandler(logger.handlers[0])
    logger.addHandler(handler)
    return logger


def get_max_rss_cpu() -> float:
    """Peak resident set size of this process in MB (0 where the resource
    module is unavailable)."""
    if not unix_like:
        return 0.0
    import resource

    return resource.getrusage(resource.RUSAGE_SELF).ru_maxrss * 1e-3


def _peak_allocated(device: int) -> int:
    # torch.cuda.max_memory_allocated flattens the allocator's whole nested
    # statistics in Python (about 0.1 ms per call); the one counter needed is
    # read from the nested form directly, with the public call as fallback
    try:
        return int(torch._C._cuda_memoryStats(device)["allocated_bytes"]["all"]["peak"])
    except (AttributeError, KeyError, TypeError, RuntimeError):
        return torch.cuda.max_memory_allocated(device)


def get_max_rss_gpu() -> float:
    """Largest peak of torch-allocated device memory over the devices, in MB
    (the per-step statistics read it after every time step)."""
    peaks = [_peak_allocated(d) for d in range(torch.cuda.device_count())]
    return max(peaks) * 1e-6


def readout_with_error(c: str, *, p_false_pos: float, p_false_neg: float) -> str:
    """One bit read out with errors (reference emu_base/utils.py:102-111):
    one uniform draw per bit, '0' reads as '1' below ``p_false_pos``, '1'
    reads as '0' below ``p_false_neg``."""
    r = random.random()
    if c == "0":
        return "1" if r < p_false_pos else c
    if c == "1":
        return "0" if r < p_false_neg else c
    return c


def apply_measurement_errors(bitstrings: Counter, *, p_false_pos: float,
                             p_false_neg: float) -> Counter:
    """Resample a bag of bitstrings with readout errors, every bit of every
    shot through ``readout_with_error`` in order (reference
    emu_base/utils.py:114-136), so a seeded ``random`` gives the same bag."""
    out: Counter = Counter()
    for bits, count in bitstrings.items():
        for _ in range(count):
            out["".join(readout_with_error(ch, p_false_pos=p_false_pos,
                                           p_false_neg=p_false_neg) for ch in bits)] += 1
    return out
