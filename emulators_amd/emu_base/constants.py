"""Device count probe, as emu_base/constants.py:1-3 of the reference."""

import torch

DEVICE_COUNT = torch.cuda.device_count()
