"""Served stream across the resident pool's size range, with a fast multipole request in the mix."""
import math

import numpy as np
import pytest

from hornsculpt_amd import bem

pytestmark = pytest.mark.gpu

RHO, C = 1.205, 343.0


# Served stream across the pool's size range: requests on resident meshes of 1 992,
# 10 512 and 58 088 triangles (the served-stream pool's smallest, middle and second
# largest), dense FP64, dense mixed and one fast multipole request on the large mesh,
# all submitted together to the serving loop, then each solved alone. Served equals
# alone: dense FP64 to 1e-10, mixed to 2e-5 (split-K summation order of the FP32
# products may differ), fast multipole to 1e-6 (its leaf moments accumulate with
# atomics; both solves stop at relative residual 1e-6).
def test_served_stream_across_sizes_with_fmm_equals_alone():
    from hornsculpt_amd import variants

    lo, hi, size = 2000.0, 120000.0, 8
    targets = [lo * (hi / lo) ** (i / (size - 1)) for i in range(size)]
    pool = variants.horn_variants(size, seed=20261204, triangles=(lo, hi), max_edge=math.inf, targets=targets)
    picks = {0: pool[0].mesh(), 3: pool[3].mesh(), 6: pool[6].mesh()}
    assert [len(m.triangles) for m in picks.values()] == [1992, 10512, 58088]
    solvers = {i: bem.HipBemSolver(m) for i, m in picks.items()}
    caps = {i: 0.999 * 2 * C / (2 * math.pi * variants.maximum_edge(m)) for i, m in picks.items()}

    def request(i, freqs, precision=bem.PRECISION_FP64, fmm=False):
        num = bem.BemNumerics(solver=bem.SOLVER_GMRES, precision=precision, gmres_restart=50,
                              gmres_maximum_iterations=2000, relative_residual_tolerance=1e-6 if fmm else 1e-8)
        if fmm:
            num.acceleration = bem.ACCELERATION_FAST_MULTIPOLE
            num.fmm_leaf_capacity = 64
            num.fmm_restart_length = 50
            num.fmm_maximum_iterations = 500
        return bem.BemRequest(medium=bem.AcousticMedium(RHO, C), frequencies=freqs,
                              excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
                              interaction_plan_frequencies=sorted(freqs + [caps[i]]), numerics=num)

    work = [(0, request(0, [300.0, 1200.0, 4000.0], bem.PRECISION_MIXED), 2e-5),
            (3, request(3, [800.0, 6000.0]), 1e-10),
            (6, request(6, [400.0], fmm=True), 1e-6),
            (0, request(0, [2500.0]), 1e-10)]
    server = bem.HipBemServer(slots=3)
    print("served stream across sizes: submitting", flush=True)
    tickets = [server.submit(solvers[i], r) for i, r, _ in work]
    served = [server.poll(t) for t in tickets]
    stats = server.stats()
    server.close()
    assert stats["failed"] == 0 and stats["max_running"] >= 2
    for (i, r, tol), s in zip(work, served):
        alone = solvers[i].solve(r)
        for fa, fs in zip(alone.frequencies, s.frequencies):
            a, b = fa.excitations[0].boundary.pressure_trace, fs.excitations[0].boundary.pressure_trace
            err = np.linalg.norm(a - b) / np.linalg.norm(a)
            print(f"mesh {len(picks[i].triangles)} {fa.frequency:.0f} Hz: served vs alone {err:.1e} "
                  f"(residual {fs.excitations[0].boundary_residual:.1e})", flush=True)
            assert err <= tol
            if r.numerics.acceleration == bem.ACCELERATION_FAST_MULTIPOLE:
                assert fs.fmm is not None and fs.fmm["m2l_links"] > 0
