"""Batched shots (GPU): every problem of a batch advanced with batched launches
equals the same problem stepped alone, bit for bit; graph replay of the same
step range reproduces itself; autotuning leaves the states untouched."""
import numpy as np
import pytest

import common as C
import twave as T

pytestmark = pytest.mark.gpu

# (mx, my, bc) of the problems: narrow and wide, SBP and periodic sides
SHAPES = [(40, 48, "FCCC"), (64, 96, "CCCC"), (33, 70, "PPPP"), (80, 37, "FFPP"), (20, 300, "CCFF"),
          (48, 264, "PPCC")]


def problem(physics, precision, cpst, mx, my, bc, seed, nt=200, coef16=0):
    kw = C.cfg_kwargs(physics=physics, mx=mx, my=my, bc=bc, precision=precision, cpst=cpst, nt=nt,
                      stencil_form=1 if physics == "acoustic" else 0, coef16=coef16)
    s = T.Solver(T.config(**kw))
    rng = np.random.default_rng(seed)
    ks = [T.C_BINV_VX, T.C_BINV_VY, T.C_KAPPA] if physics == "acoustic" else \
         [T.C_BINV_VX, T.C_BINV_VY, T.C_MU, T.C_LAMBDA, T.C_TWOMU]
    for k in ks:
        lo, hi = (1.5, 2.5) if k == T.C_LAMBDA else (0.8, 1.2)
        s.set_coef(k, rng.uniform(lo, hi, s.get_coef(k).shape))
    for f in s.fields:
        s.set_field(f, 1e-3 * rng.standard_normal(s.shape(f)))
    grid = T.GRID_MM if physics == "acoustic" else T.GRID_NM
    s.set_source_ricker(grid, mx // 2, my // 3, 25.0, 0.01)
    return s


def pairs(physics, precision, cpst, shapes=SHAPES, coef16=0):
    a = [problem(physics, precision, cpst, *sh, seed=i, coef16=coef16) for i, sh in enumerate(shapes)]
    b = [problem(physics, precision, cpst, *sh, seed=i, coef16=coef16) for i, sh in enumerate(shapes)]
    return a, b


def assert_same(a, b):
    for x, y in zip(a, b):
        for f in x.fields:
            np.testing.assert_array_equal(x.get_field(f), y.get_field(f), err_msg=T.FIELD_NAMES[f])
            np.testing.assert_array_equal(x.get_companion(f), y.get_companion(f), err_msg=T.FIELD_NAMES[f])


@pytest.mark.parametrize("precision,cpst", [("fp32", "0"), ("bf16", "K"), ("bf16", "0"), ("fp64", "6R"), ("bf16", "K8")])
def test_batch_equals_solo_acoustic(precision, cpst):
    print("case batch acoustic", precision, cpst, flush=True)
    batched, solo = pairs("acoustic", precision, cpst)
    B = T.Batch(batched)
    B.step(17)
    B.step(13)  # a second step range: new graph, continues from step 17
    for s in solo:
        s.step(30)
    assert_same(batched, solo)
    assert all(s.current_step() == 30 for s in batched)


@pytest.mark.parametrize("physics,precision,cpst", [("acoustic", "bf16", "0"), ("acoustic", "fp16", "K8"),
                                                    ("elastic", "bf16", "K8")])
def test_batch_equals_solo_coef8(physics, precision, cpst):
    """8-bit coefficients (per-row affine codes) in the batched launches."""
    print("case batch coef8", physics, precision, cpst, flush=True)
    shapes = SHAPES if physics == "acoustic" else [(40, 48, "FFPP"), (33, 70, "PPPP"), (64, 300, "FFFF")]
    batched, solo = pairs(physics, precision, cpst, shapes, coef16=2)
    B = T.Batch(batched)
    B.step(20)
    for s in solo:
        s.step(20)
    assert_same(batched, solo)


def test_batch_equals_solo_elastic():
    shapes = [(40, 48, "FFPP"), (33, 70, "PPPP"), (64, 300, "FFFF")]
    batched, solo = pairs("elastic", "fp32", "6R", shapes)
    B = T.Batch(batched)
    B.step(20)
    for s in solo:
        s.step(20)
    assert_same(batched, solo)


def test_batch_graph_replay():
    batched, _ = pairs("acoustic", "bf16", "K")
    blobs = [s.checkpoint() for s in batched]
    B = T.Batch(batched)
    B.step(12)
    first = [{f: s.get_field(f) for f in s.fields} for s in batched]
    for s, blob in zip(batched, blobs):
        s.restore(blob)
    B.step(12)  # same step range: the captured graph is replayed
    for s, ref in zip(batched, first):
        for f in s.fields:
            np.testing.assert_array_equal(s.get_field(f), ref[f])


def test_batch_autotune_keeps_state():
    batched, solo = pairs("acoustic", "fp32", "0")
    B = T.Batch(batched)
    B.step(5)
    for s in solo:
        s.step(5)
    rows = B.autotune(4)
    assert rows in (4, 8, 16, 32, 64)
    assert_same(batched, solo)
    assert all(s.current_step() == 5 for s in batched)
    B.step(5)
    for s in solo:
        s.step(5)
    assert_same(batched, solo)


def test_batch_rejects_mixed():
    a = problem("acoustic", "fp32", "0", 40, 48, "FCCC", 0)
    b = problem("acoustic", "bf16", "0", 40, 48, "FCCC", 1)
    with pytest.raises(T.TwError):
        T.Batch([a, b])


def test_batch_rejects_fused(monkeypatch):
    monkeypatch.setenv("TW_KERNEL", "3")
    a = problem("elastic", "bf16", "0", 40, 48, "FFFF", 0)
    with pytest.raises(T.TwError):
        T.Batch([a])
