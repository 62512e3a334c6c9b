"""Brent's method for the jump time of a Monte-Carlo trajectory.

Same interface and iterate sequence as the reference
(emu_base/math/brents_root_finding.py:5-150): ``BrentsRootFinder`` is driven
step by step (the caller evolves the state to each abscissa and reports the
norm gap there), ``find_root_brents`` wraps it for a plain function.  The
steps are Brent's (1973) published ones: inverse quadratic interpolation
through the last three points, the secant step when two ordinates nearly
coincide, and bisection whenever the interpolated step leaves the bracket or
shrinks too slowly.  They are kept in the reference's order, so the number of
function evaluations matches its tests.
"""

from __future__ import annotations

from typing import Callable, Optional


class BrentsRootFinder:
    def __init__(self, *, start: float, end: float, f_start: float, f_end: float,
                 epsilon: float = 1e-6):
        assert start <= end
        assert f_start * f_end < 0, "Function root needs to be between a and b"
        self.epsilon = epsilon
        # (b, fb) is the best point so far, (a, fa) the other end of the bracket
        self.a, self.fa, self.b, self.fb = start, f_start, end, f_end
        self._keep_best_in_b()
        # (c, fc) the previous best point, d the one before it
        self.c, self.fc = self.a, self.fa
        self.d = self.c
        self.bisection = True
        self.current_guess = self.b
        self.next_abscissa: Optional[float] = None

    def _keep_best_in_b(self) -> None:
        if abs(self.fa) < abs(self.fb):
            self.a, self.fa, self.b, self.fb = self.b, self.fb, self.a, self.fa

    def _interpolation_step(self) -> float:
        a, b, c, fa, fb, fc = self.a, self.b, self.c, self.fa, self.fb, self.fc
        if abs(fc - fa) < self.epsilon or abs(fc - fb) < self.epsilon:
            return fb * (b - a) / (fa - fb)  # secant through a and b
        # inverse quadratic interpolation through a, b and c
        s, r, t = fb / fa, fb / fc, fa / fc
        p = s * (t * (r - t) * (c - b) + (r - 1) * (b - a))
        q = (t - 1) * (s - 1) * (r - 1)
        return p / q

    def _must_bisect(self, dx: float) -> bool:
        tiny = abs(2 * self.epsilon * self.b)
        ab = self.a - self.b
        last = abs(self.b - self.c) if self.bisection else abs(self.c - self.d)
        return (abs(dx) >= abs(3 * ab / 4) or dx * ab < 0  # outside the bracket's 3/4
                or abs(dx) >= last / 2  # not at least halving the previous step
                or last < tiny)

    def get_next_abscissa(self) -> float:
        dx = self._interpolation_step()
        self.bisection = self._must_bisect(dx)
        if self.bisection:
            dx = (self.a - self.b) / 2
        self.next_abscissa = self.b + dx
        self.d = self.c
        self.c, self.fc = self.b, self.fb
        return self.next_abscissa

    def provide_ordinate(self, abscissa: float, ordinate: float) -> None:
        assert self.next_abscissa is not None and abscissa == self.next_abscissa, \
            "Something went wrong"
        if self.fa * ordinate < 0:  # the root lies between a and the new point
            self.b, self.fb = abscissa, ordinate
        else:
            self.a, self.fa = abscissa, ordinate
        self._keep_best_in_b()
        self.current_guess = self.b

    def is_converged(self, tolerance: float) -> bool:
        return abs(self.b - self.a) < tolerance


def find_root_brents(f: Callable[[float], float], *, start: float, end: float,
                     f_start: Optional[float] = None, f_end: Optional[float] = None,
                     tolerance: float = 1e-6, epsilon: float = 1e-6) -> float:
    """A zero of f in [start, end] (f changes sign there)."""
    finder = BrentsRootFinder(start=start, end=end,
                              f_start=f(start) if f_start is None else f_start,
                              f_end=f(end) if f_end is None else f_end, epsilon=epsilon)
    while not finder.is_converged(tolerance):
        x = finder.get_next_abscissa()
        finder.provide_ordinate(x, f(x))
    return finder.current_guess
