"""Logging and memory helpers (reference: emu_base/utils.py:14-135)."""

from __future__ import annotations

import logging
import os
import random
import sys
from collections import Counter
from pathlib import Path

import torch

unix_like = os.name != "nt"
if unix_like:
    from resource import RUSAGE_SELF, getrusage


def init_logging(log_level: int, log_file: Path | None) -> logging.Logger:
    """Configure the "emulators" logger (reference emu_base/utils.py:14-50):
    no propagation, one handler (stdout or a file opened with mode 'w'),
    plain "%(message)s" format."""
    logger = logging.getLogger("emulators")
    logger.propagate = False
    logger.setLevel(logging.DEBUG)
    handler: logging.Handler
    if log_file is None:
        handler = logging.StreamHandler(sys.stdout)
    else:
        handler = logging.FileHandler(log_file, mode="w")
    handler.setLevel(log_level)
    handler.setFormatter(logging.Formatter("%(message)s"))
    for h in list(logger.handlers):
        logger.removeHandler(h)
    logger.addHandler(handler)
    return logger


def get_max_rss_cpu() -> float:
    """Peak resident set size in MB."""
    return getrusage(RUSAGE_SELF).ru_maxrss * 1e-3 if unix_like else 0.0


def get_max_rss_gpu() -> float:
    """Peak device memory allocated by torch, in MB, over all devices."""
    return max(torch.cuda.max_memory_allocated(d) * 1e-6 for d in range(torch.cuda.device_count()))


def _readout(c: str, p_false_pos: float, p_false_neg: float) -> str:
    r = random.random()
    if c == "0" and r < p_false_pos:
        return "1"
    if c == "1" and r < p_false_neg:
        return "0"
    return c


def apply_measurement_errors(bitstrings: Counter, *, p_false_pos: float,
                             p_false_neg: float) -> Counter:
    """Resample a bag of bitstrings with readout errors (reference
    emu_base/utils.py:102-135)."""
    result: Counter = Counter()
    for bits, count in bitstrings.items():
        for _ in range(count):
            result["".join(_readout(c, p_false_pos, p_false_neg) for c in bits)] += 1
    return result
