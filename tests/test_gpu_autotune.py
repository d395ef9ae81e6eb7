"""Launch-shape autotuner: every candidate of a tuned shape gives the same operator,
and the tuner's decisions are reported.

- Regular tiles, trial panels per wavefront block (HSBEM_TILE_BLOCK = 4 / 8 / 16):
  the same pairs are integrated in another order across lanes, so the assembled
  matrices agree to the rounding of their atomic sums (far tiles accumulate in
  fixed point and are order-independent; general tiles add in floating point).
- FMM translations, stage-split target (HSBEM_FMM_SPLIT = 0 / 1024 / 2048 / 4096):
  the split only changes which workgroup computes which degrees, every output is
  formed by one thread in the same order; the operators of separate builds differ
  only by the order of the FP64 atomic sums of their leaf moments (1e-16 relative).
- An unpinned solve of a mesh no other test uses records one decision per tuned
  (tunable, key), each the fastest of its timed candidates."""
import math

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture
def unpinned(monkeypatch):
    for name in ("HSBEM_TILE_BLOCK", "HSBEM_FMM_SPLIT", "HSBEM_GEMV_COLS"):
        monkeypatch.delenv(name, raising=False)
    return monkeypatch


@pytest.mark.parametrize("precision,bar", [(bem.PRECISION_FP64, 1e-12), (bem.PRECISION_MIXED, 1e-6)])
@pytest.mark.parametrize("frequency", [500.0, 6000.0])  # far tiles present / general tiles only
def test_tile_trial_blocks_agree(unpinned, precision, bar, frequency):
    mesh = meshes.exponential_horn(0.012)
    solver = bem.HipBemSolver(mesh)
    k = 2 * math.pi * frequency / 343.0
    num = bem.BemNumerics(precision=precision)
    out = {}
    for block in ("8", "4", "16"):
        unpinned.setenv("HSBEM_TILE_BLOCK", block)
        out[block] = solver.assemble(num, k, k)
    for block in ("4", "16"):
        ea, eb = _rel(out[block][0], out["8"][0]), _rel(out[block][1], out["8"][1])
        print(f"{frequency:.0f} Hz precision {precision} block {block}: A vs block 8 {ea:.1e}, B {eb:.1e}")
        assert ea <= bar and eb <= bar


def test_fmm_split_targets_bit_identical(unpinned):
    mesh = meshes.exponential_horn(0.012)
    solver = bem.HipBemSolver(mesh)
    k = 2 * math.pi * 2000.0 / 343.0
    num = bem.BemNumerics()
    rng = np.random.default_rng(11)
    x = rng.standard_normal(len(mesh.vertices)) + 1j * rng.standard_normal(len(mesh.vertices))
    ys = {}
    for target in ("0", "1024", "2048", "4096"):
        unpinned.setenv("HSBEM_FMM_SPLIT", target)
        op = solver.fmm_operator(num, k, k, leaf_capacity=32)
        ys[target] = op.apply(x)
        print(f"split target {target}: apply {op.time_apply(3):.3f} ms")
        op.close()
    for target in ("1024", "2048", "4096"):
        err = _rel(ys[target], ys["0"])
        print(f"split target {target} vs no split: {err:.1e}")
        assert err <= 1e-13


def test_autotuner_reports_fastest_candidates(unpinned):
    mesh = meshes.exponential_horn(0.0135)
    solver = bem.HipBemSolver(mesh)
    panels = f"({len(mesh.triangles)} panels)"
    req = bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[400.0, 5000.0],
                         excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
                         numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, precision=bem.PRECISION_MIXED))
    res = solver.solve(req)
    tolerance = req.numerics.relative_residual_tolerance
    assert all(f.excitations[0].boundary_residual <= tolerance for f in res.frequencies)
    k = 2 * math.pi * 2000.0 / 343.0
    op = solver.fmm_operator(bem.BemNumerics(), k, k, leaf_capacity=32)
    op.close()
    report = bem.tuning_report()
    mine = [r for r in report if panels in r["key"]]
    for r in mine:
        print(r)
    tunables = {r["tunable"] for r in mine}
    assert {"tile_trial_block", "fmm_split_target"} <= tunables
    kinds = {r["key"].rsplit(" ", 1)[-1] for r in mine if r["tunable"] == "tile_trial_block"}
    assert {"far", "general"} <= kinds  # 400 Hz has far tiles, 5 kHz only general ones
    for r in mine:
        timings = dict((int(c), ms) for c, ms in r["timings"])
        assert len(timings) >= 3 and all(ms > 0 for ms in timings.values())
        assert r["chosen"] == min(timings, key=timings.get)
    nv = len(mesh.vertices)
    assert any(r["tunable"] == "gemv_shape" and r["key"] == f"n={nv} fp32" for r in report)
