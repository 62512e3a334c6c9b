"""Shape-preserving piecewise-cubic Hermite interpolation (PCHIP).

Used to sample the pulse (Omega, delta, phi) at time-step midpoints from
1 ns samples.  Same interpolant as the reference (emu_base/math/pchip_torch.py,
after C. Moler, "Numerical Computing with MATLAB", ch. 3):

* interior knot slopes: weighted harmonic mean of the neighbouring secants
  when they have the same sign, else 0;
* end slopes: one-sided three-point estimate, zeroed when it opposes the
  first secant, capped at 3x the first secant when the secants change sign;
* evaluation outside [x0, x_last] extends the boundary cubic.

The helper names match the reference so its unit tests apply unchanged.
"""

from __future__ import annotations

import torch


def _weighted_harmonic_mean(delta_l, delta_r, h_l, h_r):
    """(w_l + w_r) / (w_l/delta_l + w_r/delta_r) with w_l = h_l + 2 h_r,
    w_r = 2 h_l + h_r (pchip_torch.py:97-106)."""
    w_l = h_l + 2.0 * h_r
    w_r = 2.0 * h_l + h_r
    return (w_l + w_r) / (w_l / delta_l + w_r / delta_r)


def _endpoint_slope(delta_l, delta_r, h_l, h_r):
    """Three-point one-sided derivative estimate (pchip_torch.py:109-128)."""
    return ((2.0 * h_l + h_r) * delta_l - h_l * delta_r) / (h_l + h_r)


def _limit_endpoint(d_end, s_l, s_r):
    """Monotonicity limiter for the end slopes (pchip_torch.py:131-143)."""
    zero = torch.zeros_like(d_end)
    d = torch.where(d_end * s_l < 0, zero, d_end)
    cap = (s_l * s_r < 0) & (d.abs() > 3.0 * s_l.abs())
    return torch.where(cap, 3.0 * s_l, d)


def _pchip_derivatives(h: torch.Tensor, delta: torch.Tensor) -> torch.Tensor:
    """Knot slopes from interval widths and secants (pchip_torch.py:146-184)."""
    npts = h.numel() + 1
    if npts == 2:
        return delta[0].expand(2).clone()
    interior = _weighted_harmonic_mean(delta[:-1], delta[1:], h[:-1], h[1:])
    same_sign = delta[:-1] * delta[1:] > 0
    first = _limit_endpoint(_endpoint_slope(delta[0], delta[1], h[0], h[1]), delta[0], delta[1])
    last = _limit_endpoint(_endpoint_slope(delta[-1], delta[-2], h[-1], h[-2]), delta[-1],
                           delta[-2])
    mid = torch.where(same_sign, interior, torch.zeros_like(interior))
    return torch.cat([first.reshape(1), mid, last.reshape(1)])


class PCHIP1D:
    """1-D PCHIP interpolant of (x, y); ``pchip(xq)`` evaluates it."""

    def __init__(self, x: torch.Tensor, y: torch.Tensor):
        x = torch.as_tensor(x)
        if not x.is_floating_point():
            raise TypeError("x must be a floating point tensor")
        y = torch.as_tensor(y, dtype=x.dtype, device=x.device)
        if not y.is_floating_point():
            raise TypeError("y must be a floating point tensor")
        if x.ndim != 1 or y.ndim != 1:
            raise ValueError("x and y must be 1D tensors")
        if x.numel() != y.numel():
            raise ValueError("x and y must have the same length")
        if x.numel() < 2:
            raise ValueError("Need at least 2 points")
        if not torch.all(x[1:] > x[:-1]):
            raise ValueError("x must be strictly increasing")
        self.x, self.y = x, y
        h = x[1:] - x[:-1]
        delta = (y[1:] - y[:-1]) / h
        d = _pchip_derivatives(h, delta)
        # Hermite cubic on [x_i, x_i + h_i] in powers of t = x - x_i.
        c2 = (3.0 * delta - 2.0 * d[:-1] - d[1:]) / h
        c3 = (d[:-1] + d[1:] - 2.0 * delta) / (h * h)
        self._coeffs = torch.stack([y[:-1], d[:-1], c2, c3], dim=-1)

    def __call__(self, xq: torch.Tensor) -> torch.Tensor:
        xq = torch.as_tensor(xq, dtype=self.x.dtype, device=self.x.device)
        i = (torch.searchsorted(self.x, xq, right=True) - 1).clamp(0, self.x.numel() - 2)
        t = xq - self.x[i]
        c = self._coeffs[i]
        return c[..., 0] + t * (c[..., 1] + t * (c[..., 2] + t * c[..., 3]))
