"""Run a problem on the CPU restatement or on the device through the C-ABI."""
import pytest

BACKENDS = ["oracle", pytest.param("device", marks=pytest.mark.gpu)]


def make(backend, grid, params, fields):
    if backend == "oracle":
        from oracle.oracle import Oracle
        return Oracle(grid, params, fields)
    from lare3d_amd.device import Device
    print(f"\n[device case] n={grid.n} bcs={grid.bcs}", flush=True)
    return Device(grid, params, fields)


def start(sim):
    sim.boundary_conditions()
    sim.eta_calc()


def run_until(sim, t_end, max_steps=100000):
    step, t = 0, 0.0
    while t < t_end and step < max_steps:
        step += 1
        _, t = sim.step(step)
    return step, t


def close(sim):
    if hasattr(sim, "close"):
        sim.close()
