"""Batched shots (GPU): every problem of a batch advanced with batched launches
equals the same problem stepped alone, bit for bit; graph replay of the same
step range reproduces itself; autotuning leaves the states untouched."""
import numpy as np
import pytest

import common as C
import oracle as O
import twave as T

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def streaming_solo(monkeypatch):
    """Solo runs on the streaming kernels the batch runs: on grids this small
    tw_create would pick the per-point kernels, whose fields and traces are
    the same bit for bit but whose energies sum their terms in another order."""
    monkeypatch.setenv("TW_KERNEL", "2")

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


@pytest.mark.parametrize("precision,cpst", [("fp32", "0"), ("bf16", "K"), ("bf16", "0"), ("fp64", "6R"), ("bf16", "K8"),
                                            ("fp16", "K8S")])
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


def oracle_twin(physics, precision, cpst, mx, my, bc, seed, nt=200, coef16=0):
    """The CPU oracle set up like problem(): same configuration, the
    coefficients the solver decoded (8-bit codes: FP32 coefficients set to the
    decoded values), the same seeded fields and source."""
    kw = C.cfg_kwargs(physics=physics, mx=mx, my=my, bc=bc, precision=precision, cpst=cpst, nt=nt,
                      stencil_form=1 if physics == "acoustic" else 0, coef16=0 if coef16 >= 2 else coef16)
    o = O.Oracle(O.make_config(**kw))
    s = problem(physics, precision, cpst, mx, my, bc, seed, nt=nt, coef16=coef16)
    ks = [T.C_BINV_VX, T.C_BINV_VY, T.C_KAPPA] if physics == "acoustic" else \
         [T.C_BINV_VX, T.C_BINV_VY, T.C_MU, T.C_LAMBDA, T.C_TWOMU]
    for k in ks:
        o.set_coef(k, s.get_coef(k))
    for f in s.fields:
        o.set_field(f, s.get_field(f))
    grid = T.GRID_MM if physics == "acoustic" else T.GRID_NM
    o.set_source_ricker(grid, mx // 2, my // 3, 25.0, 0.01)
    s.close()
    return o


@pytest.mark.parametrize("physics,precision,cpst,coef16", [("acoustic", "bf16", "K8", 1), ("acoustic", "bf16", "0", 2),
                                                           ("elastic", "bf16", "K8", 2), ("acoustic", "fp32", "6R", 0),
                                                           ("elastic", "fp16", "K8S", 2), ("acoustic", "bf16", "K8S", 1),
                                                           ("elastic", "bf16", "K8S", 3)])
def test_batch_equals_oracle(physics, precision, cpst, coef16):
    """Batched launches against the CPU oracle directly (not only against the
    same problems stepped alone): every field and companion bitwise after a
    graph-captured step range and a second range."""
    print("case batch vs oracle", physics, precision, cpst, coef16, flush=True)
    shapes = SHAPES if physics == "acoustic" else [(40, 48, "FFPP"), (33, 70, "PPPP"), (64, 300, "FFFF")]
    batched = [problem(physics, precision, cpst, *sh, seed=i, coef16=coef16) for i, sh in enumerate(shapes)]
    twins = [oracle_twin(physics, precision, cpst, *sh, seed=i, coef16=coef16) for i, sh in enumerate(shapes)]
    B = T.Batch(batched)
    B.step(25)
    B.step(15)
    for o in twins:
        o.step(40)
    for x, o in zip(batched, twins):
        C.compare_state(x, o, companions=True)
    B.close()


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


def test_batch_graph_rebuilt_after_new_source():
    """Reset, set the next shot's source, step again from the same step: the
    batch must not replay the graph captured with the previous source."""
    batched, solo = pairs("acoustic", "bf16", "K8")
    blobs = [s.checkpoint() for s in batched]
    B = T.Batch(batched)
    B.step(12)
    for s, blob in zip(batched, blobs):
        s.restore(blob)
        s.set_source_ricker(T.GRID_MM, s.shape(T.P)[0] // 3, s.shape(T.P)[1] // 2, 18.0, 0.02)
    for s in solo:
        s.set_source_ricker(T.GRID_MM, s.shape(T.P)[0] // 3, s.shape(T.P)[1] // 2, 18.0, 0.02)
    B.step(12)
    for s in solo:
        s.step(12)
    assert_same(batched, solo)
    # and new coefficients between replays
    for s, blob in zip(batched, blobs):
        s.restore(blob)
    for x, y in zip(batched, solo):
        k = T.C_KAPPA
        v = 1.1 * x.get_coef(k)
        x.set_coef(k, v)
        y.restore(x.checkpoint())
        y.set_coef(k, v)
    B.step(12)
    for s in solo:
        s.step(12)
    assert_same(batched, solo)


def with_receivers(solvers, seed):
    for i, s in enumerate(solvers):
        rng = np.random.default_rng(seed + i)
        fs, xs, ys = [], [], []
        for f in s.fields:
            r, c = s.shape(f)
            for _ in range(3):
                fs.append(f)
                xs.append(int(rng.integers(0, r)))
                ys.append(int(rng.integers(0, c)))
        s.set_receivers(fs, xs, ys)
    return solvers


@pytest.mark.parametrize("physics,precision,cpst", [("acoustic", "bf16", "K8"), ("acoustic", "fp64", "0"),
                                                    ("elastic", "fp32", "6R"), ("elastic", "fp16", "K8S")])
def test_batch_records_receivers_and_energy(physics, precision, cpst):
    """Batched steps record every solver's receiver samples and energies as
    tw_step does: traces bitwise, energies bitwise at the same tile height;
    graph replay from a restored state records the same again."""
    print("case batch records", physics, precision, cpst, flush=True)
    shapes = SHAPES if physics == "acoustic" else [(40, 48, "FFPP"), (33, 70, "PPPP"), (64, 300, "FFFF")]
    batched, solo = pairs(physics, precision, cpst, shapes)
    with_receivers(batched, 50)
    with_receivers(solo, 50)
    blobs = [s.checkpoint() for s in batched]
    B = T.Batch(batched)
    B.step(17)
    B.step(13)
    for s in solo:
        s.step(17)
        s.step(13)
    assert_same(batched, solo)
    for x, y in zip(batched, solo):
        np.testing.assert_array_equal(x.receivers(), y.receivers())
        for a, b in zip(x.energy(), y.energy()):
            np.testing.assert_array_equal(a, b)
        assert np.abs(x.receivers()).max() > 0 and np.all(np.isfinite(x.energy()[0]))
    first = [(x.receivers(), x.energy()) for x in batched]
    for x, blob in zip(batched, blobs):
        x.restore(blob)
    B.step(17)  # replays the first graph (same step range, same solvers)
    B.step(13)
    for x, (r, e) in zip(batched, first):
        np.testing.assert_array_equal(x.receivers(), r)
        for a, b in zip(x.energy(), e):
            np.testing.assert_array_equal(a, b)
    B.close()


def test_batch_autotune_then_solo_steps_record_energy():
    """After the autotuner picks a tile height for the batch, each solver's
    own steps keep their partial-sum layout: a batched range and then solo
    steps on the same solvers give the energies and traces of solo runs (to
    rounding for the batched range, whose tile height may differ)."""
    batched, solo = pairs("acoustic", "fp32", "0")
    with_receivers(batched, 60)
    with_receivers(solo, 60)
    B = T.Batch(batched)
    B.autotune(4)
    B.step(10)
    for s in batched:
        s.step(6)
    for s in solo:
        s.step(16)
    assert_same(batched, solo)
    for x, y in zip(batched, solo):
        np.testing.assert_array_equal(x.receivers(), y.receivers())
        for a, b in zip(x.energy(), y.energy()):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=0)
            np.testing.assert_array_equal(a[11:], b[11:])
    B.close()


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
