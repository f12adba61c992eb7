"""Kernel functors beyond counting: the base class the traversal dispatches on.

The SPH density / pressure-force and pair-energy functors and their adjoints subclass
KernelFunctor and implement `run_grid` (grid traversal) and `run_lists`
(precomputed-list traversal)."""

from __future__ import annotations


class KernelFunctor:
    """Base of the interaction functors that run as dedicated HIP kernels."""

    def run_grid(self, nhs, x, y, points):  # pragma: no cover - abstract
        raise NotImplementedError

    def run_lists(self, nhs, x, y, points):  # pragma: no cover - abstract
        raise NotImplementedError
