"""Device-resident drivers with the shape of magex's GPU classes (magex/cuda_backend.py:338-418).

``HipLLG(system, alpha, dt, temperature, seed, integrator, precession)`` keeps the spins on the
device between calls: ``step_device(e_dev)`` advances a device tensor by one step and ``run(e, steps,
callback, every)`` copies in, steps, and copies out (calling ``callback(n, e_host)`` every ``every``
steps).  ``CudaLLG`` is the same class under magex's name, so ``from magex.cuda_backend import CudaLLG``
becomes ``from magex_amd.hip_backend import CudaLLG``.  Same equations, units and (counter-based)
noise as ``LLGSolver``; the kernels are libmxllg's.  ``HipMC`` / ``CudaMC`` is magex's GPU Metropolis
class (``sweep_device(e_dev)``, ``acceptance()``) over the library's colour sweep.
"""
from __future__ import annotations

import numpy as np

from .llg import LLGSolver
from .montecarlo import MetropolisMC


class HipLLG:
    """Stochastic Heun / Depondt LLG on the GPU (magex/cuda_backend.py:338-377)."""

    def __init__(self, system, alpha=0.1, dt=1e-3, temperature=0.0, seed=1, integrator="heun", precession=True,
                 precision="fp64"):
        self.sol = LLGSolver(system, alpha=alpha, dt=dt, temperature=temperature, seed=seed, precession=precession,
                             integrator=integrator, precision=precision)
        self.sys = system
        self.dt = float(dt)
        self.T = float(temperature)

    @property
    def time(self):
        return self.sol.time

    def step_device(self, e_dev, nsteps=1):
        """Advance the device tensor e_dev (N, 3) in place."""
        self.sol.T = self.T
        self.sol.step_device(e_dev, nsteps)
        return e_dev

    def run(self, e, steps, callback=None, every=0):
        """steps steps from host spins e; returns host spins (magex/cuda_backend.py:371-377)."""
        e_dev = self.sol._to_device(np.ascontiguousarray(e, float))
        if e_dev is e:
            e_dev = e_dev.clone()
        done = 0
        while done < steps:
            n = steps - done
            if callback and every:
                n = min(n, every - done % every)
            self.sol.T = self.T
            self.sol.steps_checked(e_dev, n)
            done += n
            if callback and every and done % every == 0:
                callback(done, e_dev.cpu().numpy().astype(float))
        return e_dev.cpu().numpy().astype(float)


CudaLLG = HipLLG


class HipMC:
    """Metropolis / heat-bath colour sweeps on the GPU (magex/cuda_backend.py:380-418): no
    over-relaxation pass; heat-bath moves fall back to gaussian ones where some site has T = 0, as
    magex's class does.  The random numbers are the library's counter-based draws."""

    def __init__(self, system, temperature, seed=1, sigma=0.6, move="gaussian", precision="fp64"):
        from .structure import site_temperature
        T = np.broadcast_to(site_temperature(system, float(temperature)), (system.N,))
        if move == "heatbath" and not np.all(np.asarray(T, float) > 0):
            move = "gaussian"
        self.sys = system
        self.sigma = float(sigma)
        self.mc = MetropolisMC(system, temperature, seed=seed, sigma=sigma, overrelax=0, move=move,
                               precision=precision)

    @property
    def tried(self):
        return self.mc.tried

    def sweep_device(self, e_dev):
        """One sweep over every colour of the device tensor e_dev (N, 3), in place."""
        self.mc.sweep_device(e_dev, 1)
        return e_dev

    def acceptance(self):
        """Accepted fraction since the last call; resets the counters (magex/cuda_backend.py:414-418)."""
        return self.mc.acceptance(reset=True)


CudaMC = HipMC
