"""The batched-meshes workload: independent seeded Mooney-Rivlin unit-cube
meshes of uneven size (n^3 triquadratic hexahedra, n ~ U{n_min..n_max}),
materials c10 ~ U[1,5], c01 ~ U[0,10], uniaxial stretch ~ U[5%,20%] over 5
load steps, interior nodes jittered by 0.1h, all drawn from one master seed
with a per-mesh seed derived from it."""
import time

import numpy as np

from .equations_set import MooneyRivlinBatch
from .mesh import RegularMesh


def draw_cases(n_meshes=256, seed=2024, n_min=4, n_max=16):
    rng = np.random.default_rng(seed)
    return [dict(n=int(rng.integers(n_min, n_max + 1)), c10=float(rng.uniform(1, 5)), c01=float(rng.uniform(0, 10)),
                 stretch=float(rng.uniform(0.05, 0.20)), seed=int(rng.integers(0, 2 ** 31)))
            for _ in range(n_meshes)]


def build(cases, jitter=0.1):
    """(meshes, batch equations set at z = X, p = 0 with the uniaxial
    conditions, setup seconds)."""
    t0 = time.perf_counter()
    meshes = [RegularMesh.generate(2, c["n"], jitter=jitter, seed=c["seed"]) for c in cases]
    batch = MooneyRivlinBatch(meshes, [c["c10"] for c in cases], [c["c01"] for c in cases])
    batch.set_dofs(batch.undeformed_dofs())
    batch.fixed_dofs = batch.uniaxial_boundary_conditions([1.0 + c["stretch"] for c in cases])
    return meshes, batch, time.perf_counter() - t0
