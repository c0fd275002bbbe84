"""Reference-format front-end on the GPU: a problem given as InputFile.txt +
SrcRcv.txt (+ model files) runs through twave.frontend.run_case, which writes
the reference's Signal_*/Energy_* text files; every number in them equals the
oracle's run of the same shot (receiver fields in the reference's per-grid
order, E = E_k + (E_p0 + E_p1)/2), to the 16 significant digits the format
prints."""
import os

import numpy as np
import pytest

import common as C
import oracle as O
import twave as T
import twave.frontend as FE

pytestmark = pytest.mark.gpu

# 48 (physical X, data y) x 40 (physical Y, data x) cells, dx = 1/125
INPUT = """model_24_20     // model on a twice coarser grid
dxM_2_125
dxS_1_125
grids_48_40
Nt_{nt}
MAXdt_1.6e-04
CFL_0.625
frequency_25.0
delay_0.04
"""
# locations in physical X, Y (num/den of length units); grids physical X first
# (units of 1/250: physical X spans 0..96, Y 0..80; an M index is (num-1)/2,
# an N index num/2)
SRCRCV = """Src_3_MN_41_250_20_250
Rcv_3_1_MN_61_250_20_250
Rcv_3_2_NN_40_250_40_250
Rcv_3_4_MM_31_250_61_250
Rcv_3_3_NM_20_250_71_250
Src_7_MM_51_250_21_250
Rcv_7_1_NN_30_250_50_250
Rcv_7_5_MM_11_250_31_250
"""
FIELDS = {0: [T.SXY], 1: [T.VY], 2: [T.VX], 3: [T.SXX, T.SYY]}


def oracle_shot(cfg, inp, entries, n, k, model):
    oc = O.make_config(physics=0, mx=cfg.mx, my=cfg.my, dx=cfg.dx, num_dx=cfg.num_dx, den_dx=cfg.den_dx,
                       dt=cfg.dt, bc=cfg.bc.decode(), precision=cfg.precision, nt_max=cfg.nt_max)
    o = O.Oracle(oc)
    o.set_model(*model, inp.num_dx_model, inp.den_dx_model)
    src = [entries[i] for i in range(n) if entries[i].kind == 0 and entries[i].src_index == k][0]
    g = FE.GRID_OF[src.grid]
    ix, iy = T.locate(cfg, g, src.num[0], src.den[0], src.num[1], src.den[1])
    o.set_source_ricker(g, ix, iy, inp.central_f, inp.time_delay)
    fs, xs, ys, rcv = [], [], [], []
    for i in range(n):
        r = entries[i]
        if r.kind != 1 or r.src_index != k:
            continue
        rg = FE.GRID_OF[r.grid]
        rx, ry = T.locate(cfg, rg, r.num[0], r.den[0], r.num[1], r.den[1])
        rcv.append((r.rcv_index, r.grid, rx, ry, len(FIELDS[rg])))
        for f in FIELDS[rg]:
            fs.append(f)
            xs.append(rx)
            ys.append(ry)
    o.set_receivers(fs, xs, ys)
    o.step(cfg.nt_max)
    return (g, ix, iy, src.grid), rcv, o.receivers(), o.energy()


def tag(grid, ix, iy):
    return "%s%s_%d_%d" % (chr(grid[1]), chr(grid[0]), iy, ix)


@pytest.mark.parametrize("with_models", [False, True])
def test_run_case_files_match_oracle(tmp_path, with_models):
    print("case frontend run_case models=%s" % with_models, flush=True)
    nt = 60
    (tmp_path / "InputFile.txt").write_text(INPUT.format(nt=nt))
    (tmp_path / "SrcRcv.txt").write_text(SRCRCV)
    inp = FE.read_input_file(str(tmp_path / "InputFile.txt"))
    mx, my = inp.model_m[0], inp.model_m[1]
    if with_models:
        rng = np.random.default_rng(3)
        rho = rng.uniform(0.9, 1.1, (mx + 1, my + 1))
        vp = rng.uniform(1.8, 2.2, (mx + 1, my + 1))
        vs = vp / 1.8
        for name, a in (("rho", rho), ("vp", vp), ("vs", vs)):
            a[:, -1] = a[:, 0]          # periodic y needs equal first/last columns
            (tmp_path / (name + ".bin")).write_bytes(a.astype("<f8").tobytes())
        model = (rho, vp, vs)
        mdir = str(tmp_path)
    else:
        model = tuple(np.full((mx + 1, my + 1), v) for v in (1.0, 2.0, 1.0))
        mdir = None
    out = FE.run_case(str(tmp_path / "InputFile.txt"), str(tmp_path / "SrcRcv.txt"), str(tmp_path / "result"),
                      model_dir=mdir)
    assert sorted(out) == [3, 7]
    cfg = FE.input_to_config(inp)
    cfg.bc = b"FFPP"
    entries, n = FE.read_srcrcv_file(str(tmp_path / "SrcRcv.txt"))
    folder = out[3]
    assert folder == FE.result_folder(inp, cfg, str(tmp_path / "result"))
    for k in (3, 7):
        (g, ix, iy, sgrid), rcv, traces, (ek, ep0, ep1) = oracle_shot(cfg, inp, entries, n, k, model)
        row = 0
        for (j, rgrid, rx, ry, nf) in rcv:
            name = "Signal_src_%d_%s_rcv_%d_%s.txt" % (k, tag(sgrid, ix, iy), j, tag(rgrid, rx, ry))
            data = np.loadtxt(os.path.join(folder, name))
            assert data.shape == (nt, 1 + nf), name
            np.testing.assert_array_equal(data[:, 0], np.arange(nt))
            np.testing.assert_allclose(data[:, 1:].T, traces[row:row + nf], rtol=1e-15, atol=1e-300, err_msg=name)
            row += nf
        e = np.loadtxt(os.path.join(folder, "Energy_src_%d_%s.txt" % (k, tag(sgrid, ix, iy))))
        np.testing.assert_allclose(e, ek + (ep0 + ep1) / 2.0, rtol=1e-12, atol=0)
        assert np.abs(e).max() > 0


@pytest.mark.parametrize("physics", ["elastic", "acoustic"])
def test_write_snapshots(tmp_path, physics):
    """Snapshot_<type>_<name>_<i>.txt of every field in the reference's grid
    order and names (nvcc_device_FDM_2D.cu:788-819), values = stored state
    times source_scaling, "% 16.15e", row-major."""
    s, o = C.make_pair(physics=physics, mx=24, my=30, bc="FFPP", precision="fp32", cpst="6R", nt=12,
                       source_scaling=2.0, stencil_form=1 if physics == "acoustic" else 0)
    C.random_fields(s, o, 3)
    s.step(12)
    o.step(12)
    FE.write_snapshots(s, str(tmp_path))
    if physics == "elastic":
        names = {"NN_Sxy_0": T.SXY, "NM_Vx_0": T.VY, "MN_Vy_0": T.VX, "MM_SMM_0": T.SXX, "MM_SMM_1": T.SYY}
    else:
        names = {"NM_Vx_0": T.VY, "MN_Vy_0": T.VX, "MM_SMM_0": T.P}
    files = sorted(p.name for p in tmp_path.iterdir())
    assert files == sorted("Snapshot_%s.txt" % k for k in names), files
    for k, f in names.items():
        got = np.loadtxt(str(tmp_path / ("Snapshot_%s.txt" % k)))
        want = np.array([float("% 16.15e" % v) for v in (2.0 * o.get_field(f)).ravel()])
        np.testing.assert_array_equal(got, want, err_msg=k)
    s.close()


def test_write_results_requires_the_shots_receivers(tmp_path):
    """Results are written only when the solver records exactly the receivers
    of the shot named (count, fields and indices); otherwise TW_ESTATE, never
    a write past the records or traces under other receivers' names."""
    (tmp_path / "InputFile.txt").write_text(INPUT.format(nt=10))
    inp = FE.read_input_file(str(tmp_path / "InputFile.txt"))
    cfg = FE.input_to_config(inp)
    s = T.Solver(cfg)
    (tmp_path / "a.txt").write_text("Src_1_MN_41_250_20_250\nRcv_1_1_MN_61_250_20_250\n"
                                    "Src_2_MN_81_250_20_250\nRcv_2_1_MN_61_250_20_250\nRcv_2_2_MN_71_250_30_250\n")
    entries, n = FE.read_srcrcv_file(str(tmp_path / "a.txt"))
    FE.setup_shot(s, inp, entries, n, 1)
    f, x, y = s.get_receivers()
    assert len(f) == s.nrcv > 0
    s.step(5)
    FE.write_results(s, str(tmp_path), entries, n, 1)
    with pytest.raises(T.TwError) as e:
        FE.write_results(s, str(tmp_path), entries, n, 2)  # shot 2 has more receivers
    assert e.value.code == 6  # TW_ESTATE
    s.set_receivers(list(f), list(x + 1), list(y))  # same count, other points
    s.step(1)
    with pytest.raises(T.TwError) as e:
        FE.write_results(s, str(tmp_path), entries, n, 1)
    assert e.value.code == 6  # TW_ESTATE
    s.close()


def test_setup_shot_rejects(tmp_path):
    (tmp_path / "InputFile.txt").write_text(INPUT.format(nt=10))
    inp = FE.read_input_file(str(tmp_path / "InputFile.txt"))
    cfg = FE.input_to_config(inp)
    s = T.Solver(cfg)
    (tmp_path / "a.txt").write_text("Src_1_MN_41_250_20_250\n")          # no receivers
    entries, n = FE.read_srcrcv_file(str(tmp_path / "a.txt"))
    with pytest.raises(T.TwError):
        FE.setup_shot(s, inp, entries, n, 1)
    with pytest.raises(T.TwError):
        FE.setup_shot(s, inp, entries, n, 2)                               # not in the list
    (tmp_path / "b.txt").write_text("Src_1_MN_41_250_0_250\nRcv_1_1_MN_61_250_20_250\n")     # x node boundary
    entries, n = FE.read_srcrcv_file(str(tmp_path / "b.txt"))
    with pytest.raises(T.TwError):
        FE.setup_shot(s, inp, entries, n, 1)
    (tmp_path / "c.txt").write_text("Src_1_MN_41_250_20_250\nRcv_1_1_MN_40_250_20_250\n")    # not a grid point
    entries, n = FE.read_srcrcv_file(str(tmp_path / "c.txt"))
    with pytest.raises(T.TwError):
        FE.setup_shot(s, inp, entries, n, 1)
    s.close()
