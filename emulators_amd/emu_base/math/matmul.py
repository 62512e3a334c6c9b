"""A 2x2 single-qubit operator applied along axis 1 of a batched tensor
(reference emu_base/math/matmul.py: matmul_2x2_with_batched, which the
reference's Lindbladian uses on density matrices viewed as (left, 2, right))."""

from __future__ import annotations

import torch


def matmul_2x2_with_batched(left: torch.Tensor, right: torch.Tensor) -> torch.Tensor:
    """left (2, 2) times every (2, k) slice of right (b, 2, k): the batched
    product left @ right, in right's dtype and on right's device."""
    if tuple(left.shape) != (2, 2) or right.dim() != 3 or right.shape[1] != 2:
        raise ValueError("matmul_2x2_with_batched: left must be 2 x 2 and right (b, 2, k)")
    op = left.to(dtype=right.dtype, device=right.device)
    # out[:, i, :] = op[i, 0] right[:, 0, :] + op[i, 1] right[:, 1, :]
    r0, r1 = right[:, 0, :], right[:, 1, :]
    return torch.stack((op[0, 0] * r0 + op[0, 1] * r1, op[1, 0] * r0 + op[1, 1] * r1), dim=1)
