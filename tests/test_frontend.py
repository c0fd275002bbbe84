"""Reference-format front-end, host side (no GPU): InputFile.txt and
command-line options, SrcRcv.txt, raw model files, the result folder name.

The input texts below follow the reference's formats
(cuda/src/cpp/input_processing.cpp:11-137, cuda/src/cpp/inverse.cpp:133-348);
the known answers are the reference driver's own: the shipped 600^2 case
(cuda/data/homogeneous) puts its source Src_1_MN_401_250_20_250 at data
(10, 200) on the (N, M) grid and its receiver at (10, 400), and runs with
dt = MAXdt * CFL = 1e-4.
"""
import ctypes
import struct

import numpy as np
import pytest

import twave as T
import twave.frontend as FE

INPUT = """model_600_600                   // Nx_model; Ny_model; Nz_model

dxM_1_125                       // model grid spacing (numerator_denominator);
\t\t\t\t        // continuation lines of a comment are ignored

dxS_1_125                       // simulation grid spacing (num and den)

grids_600_600                     // number of M-grid points

#bdry_PPFF                   \t// boundary types (commented out)

Nt_60000                        // total time steps

MAXdt_1.6e-04\t\t\t\t    // maximally allowed dt

CFL_0.625                       // CFL constant

frequency_5.0                   // central frequency

delay_0.25                      // time delay in the source signal

// If the simulation is in 2D, the last token in model and grids needs to be zero.
"""

SRCRCV = """# comment lines start with '#'
#
Src_1_MN_401_250_20_250
Rcv_1_1_MN_801_250_20_250
"""


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_input_file_defaults_and_values(tmp_path):
    inp = FE.read_input_file(write(tmp_path, "InputFile.txt", INPUT))
    assert list(inp.model_m) == [600, 600] and list(inp.grid_m) == [600, 600]
    assert (inp.num_dx_model, inp.den_dx_model, inp.num_dx, inp.den_dx) == (1, 125, 1, 125)
    assert inp.nt == 60000
    assert inp.dt_max == 1.6e-4 and inp.cfl == 0.625
    assert inp.central_f == 5.0 and inp.time_delay == 0.25
    assert inp.bc == b"FFFF"            # the bdry line is commented out: reference default
    assert inp.energy == b"Y" and inp.medium == b"homogeneous" and inp.device == -1
    cfg = FE.input_to_config(inp)
    assert (cfg.mx, cfg.my) == (600, 600)
    assert cfg.dx == 1.0 / 125 and (cfg.num_dx, cfg.den_dx) == (1, 125)
    assert cfg.dt == 1.6e-4 * 0.625
    assert cfg.nt_max == 60000 and cfg.record_energy == 1 and cfg.bc == b"FFFF"


def test_axes_and_boundary_letters(tmp_path):
    """Files give physical X first; physical X is data axis 1 (y)."""
    txt = INPUT.replace("grids_600_600", "grids_300_200").replace("model_600_600", "model_150_100") \
               .replace("dxM_1_125", "dxM_2_125").replace("#bdry_PPFF", "bdry_PPFC")
    inp = FE.read_input_file(write(tmp_path, "in.txt", txt))
    assert list(inp.grid_m) == [200, 300] and list(inp.model_m) == [100, 150]
    assert inp.bc == b"FCPP"            # X sides PP -> data y; Y sides FC -> data x
    cfg = FE.input_to_config(inp)
    assert (cfg.mx, cfg.my) == (200, 300)


def test_options_override_file(tmp_path):
    inp = FE.read_input_file(write(tmp_path, "in.txt", INPUT),
                             {"Nt": "200", "B_type": "ffpp", "CFL": "0.5", "f": "12", "t0": "0.1",
                              "medium": "layered", "energy": "n", "device": "0", "dt_max": "2e-4"})
    assert inp.nt == 200 and inp.bc == b"PPFF" and inp.cfl == 0.5
    assert inp.central_f == 12.0 and inp.time_delay == 0.1 and inp.medium == b"layered"
    assert inp.energy == b"N" and inp.device == 0 and inp.dt_max == 2e-4
    cfg = FE.input_to_config(inp)
    assert cfg.record_energy == 0 and cfg.dt == 2e-4 * 0.5
    for k, v in (("Nt", "x"), ("CFL", "1.0abc"), ("B_type", "FX"), ("nope", "1")):
        with pytest.raises(T.TwError):
            FE.read_input_file(write(tmp_path, "in.txt", INPUT), {k: v})


def test_config_checks(tmp_path):
    # model extent 300 * 1/125 != 600 * 1/125 (input_processing.cpp:206-210)
    bad = INPUT.replace("model_600_600", "model_300_600")
    with pytest.raises(T.TwError):
        FE.input_to_config(FE.read_input_file(write(tmp_path, "a.txt", bad)))
    ok = bad.replace("dxM_1_125", "dxM_2_125").replace("model_300_600", "model_300_300")
    FE.input_to_config(FE.read_input_file(write(tmp_path, "b.txt", ok)))
    with pytest.raises(T.TwError):
        FE.input_to_config(FE.read_input_file(write(tmp_path, "c.txt", INPUT), {"energy": "X"}))
    with pytest.raises(T.TwError):
        FE.read_input_file(write(tmp_path, "d.txt", INPUT.replace("Nt_60000", "Nt_6e4")))
    with pytest.raises(T.TwError):
        FE.read_input_file(str(tmp_path / "missing.txt"))


def test_srcrcv_shipped_case(tmp_path):
    entries, n = FE.read_srcrcv_file(write(tmp_path, "SrcRcv.txt", SRCRCV))
    assert n == 2
    s, r = entries[0], entries[1]
    assert (s.kind, s.src_index, s.rcv_index, s.grid) == (0, 1, -1, b"NM")
    assert (list(s.num), list(s.den)) == ([20, 401], [250, 250])
    assert (r.kind, r.src_index, r.rcv_index, r.grid) == (1, 1, 1, b"NM")
    inp = FE.read_input_file(write(tmp_path, "in.txt", INPUT))
    cfg = FE.input_to_config(inp)
    g = FE.GRID_OF[s.grid]
    assert T.locate(cfg, g, s.num[0], s.den[0], s.num[1], s.den[1]) == (10, 200)
    assert T.locate(cfg, g, r.num[0], r.den[0], r.num[1], r.den[1]) == (10, 400)


@pytest.mark.parametrize("text", [
    "Rcv_1_1_MN_801_250_20_250\nSrc_1_MN_401_250_20_250\n",                          # receiver before source
    "Src_1_MN_401_250_20_250\nSrc_2_MN_401_250_20_250\n",                            # duplicate location
    "Src_1_MN_401_250_20_250\nSrc_1_NN_400_250_20_250\n",                            # duplicate index
    "Src_1_MN_401_250_20_250\nRcv_1_1_MN_801_250_20_250\nRcv_1_2_MN_801_250_20_250\n",   # rcv location
    "Src_1_MN_401_250_20_250\nRcv_1_1_MN_801_250_20_250\nRcv_1_1_NN_800_250_20_250\n",   # rcv index
    "Src_1_MX_401_250_20_250\n",                                                     # grid letter
    "Src_1_MN_401_250_20\n",                                                         # short line
])
def test_srcrcv_rejects(tmp_path, text):
    with pytest.raises(T.TwError):
        FE.read_srcrcv_file(write(tmp_path, "s.txt", text))


def test_srcrcv_capacity(tmp_path):
    p = write(tmp_path, "s.txt", SRCRCV + "Rcv_1_2_MM_801_250_21_250\n")
    n = ctypes.c_int32()
    arr = (T.SrcRcvEntry * 2)()
    e = T.lib().tw_read_srcrcv_file(p.encode(), arr, 2, ctypes.byref(n))
    assert e == 2 and n.value == 3      # TW_ERANGE, count needed
    entries, n = FE.read_srcrcv_file(p)
    assert n == 3 and entries[2].grid == b"MM" and entries[2].rcv_index == 2


def test_model_file(tmp_path):
    mx, my = 6, 9
    a = np.random.default_rng(0).uniform(1, 2, (mx + 1, my + 1))
    p = tmp_path / "rho.bin"
    p.write_bytes(a.astype("<f8").tobytes())
    np.testing.assert_array_equal(FE.read_model_file(str(p), mx, my), a)
    with pytest.raises(T.TwError):
        FE.read_model_file(str(p), mx, my + 1)              # size check
    with pytest.raises(T.TwError):
        FE.read_model_file(str(p), mx, my, "FFPP")          # periodic y: first/last columns differ
    a[:, -1] = a[:, 0]
    p.write_bytes(a.astype("<f8").tobytes())
    FE.read_model_file(str(p), mx, my, "FFPP")


def test_result_folder(tmp_path):
    inp = FE.read_input_file(write(tmp_path, "in.txt", INPUT))
    cfg = FE.input_to_config(inp, precision="fp32", cpst="6R")
    assert FE.result_folder(inp, cfg, "result") == \
        "result/elastic/homogeneous/device/grid_size_1_125/Nt_60000/fp32/dt_1p000000000000000e-04/6_R"
    cfg = FE.input_to_config(inp, precision="bf16", cpst="K")
    assert FE.result_folder(inp, cfg, "r").endswith("/bf16/dt_1p000000000000000e-04/K")
    cfg8 = FE.input_to_config(inp, precision="bf16", cpst="K8")
    assert FE.result_folder(inp, cfg8, "r").endswith("/bf16/dt_1p000000000000000e-04/K8")
    # the reference's dt folder: "%16.15e" with the decimal point replaced
    assert "dt_" + ("%16.15e" % cfg.dt).replace(".", "p", 1) in FE.result_folder(inp, cfg, "r")
    small = ctypes.create_string_buffer(8)
    assert T.lib().tw_result_folder(ctypes.byref(inp), ctypes.byref(cfg), b"r", small, 8) == 2
