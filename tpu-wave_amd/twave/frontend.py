"""Reference-format front-end: the reference's problem files, source/receiver
lists, model files and text results, over the tw_* front-end entry points.

`run_case` is the reference's elastic driver flow
(cuda/src/cpp/main_host_FDM_2D.tpp:11-280, device twin
cuda/src/cuda_main/nvcc_device_FDM_2D.cu): read InputFile.txt, apply
command-line options, read SrcRcv.txt, set the material (the driver's
constant rho = 1, vp = 2, vs = 1, or model files), and for each source reset
the fields, run Nt steps on the GPU and write Signal_*.txt / Energy_*.txt into
the reference's result folder layout.
"""
import ctypes
import os

import numpy as np

from . import (ELASTIC, FP32, CPST, PRECISION, Config, Input, Solver, SrcRcvEntry, TwError, lib)

GRID_OF = {b"NN": 0, b"MN": 1, b"NM": 2, b"MM": 3}


def _err(e):
    if e:
        raise TwError(e, lib().tw_global_error().decode())


def read_input_file(path, options=None):
    """tw_input from an InputFile.txt, then command-line style options
    ({"Nt": "100", "B_type": "FFPP", ...}) applied over it."""
    inp = Input()
    lib().tw_input_default(ctypes.byref(inp))
    _err(lib().tw_read_input_file(os.fsencode(path), ctypes.byref(inp)))
    for k, v in (options or {}).items():
        _err(lib().tw_input_option(ctypes.byref(inp), k.encode(), str(v).encode()))
    return inp


def input_to_config(inp, physics=ELASTIC, precision=FP32, cpst="0"):
    c = Config()
    p = PRECISION[precision] if isinstance(precision, str) else precision
    k = CPST[cpst] if isinstance(cpst, str) else cpst
    _err(lib().tw_input_to_config(ctypes.byref(inp), physics, p, k, ctypes.byref(c)))
    return c


def read_srcrcv_file(path):
    """Entries of a SrcRcv.txt in file order (SrcRcvEntry array)."""
    n = ctypes.c_int32()
    _err(lib().tw_read_srcrcv_file(os.fsencode(path), None, 0, ctypes.byref(n)))
    arr = (SrcRcvEntry * max(n.value, 1))()
    _err(lib().tw_read_srcrcv_file(os.fsencode(path), arr, n.value, ctypes.byref(n)))
    return arr, n.value


def read_model_file(path, mx_model, my_model, bc=b"FFFF"):
    out = np.zeros((mx_model + 1, my_model + 1))
    bc = bc.encode() if isinstance(bc, str) else bc
    _err(lib().tw_read_model_file(os.fsencode(path), mx_model, my_model, bc,
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def result_folder(inp, cfg, root):
    buf = ctypes.create_string_buffer(4096)
    _err(lib().tw_result_folder(ctypes.byref(inp), ctypes.byref(cfg), os.fsencode(root), buf, len(buf)))
    return buf.value.decode()


def setup_shot(solver, inp, entries, n, src_index):
    solver._chk(lib().tw_setup_shot(solver.h, ctypes.byref(inp), entries, n, src_index))
    solver.get_receivers()  # keeps solver.nrcv in step with the registration


def write_results(solver, folder, entries, n, src_index):
    solver._chk(lib().tw_write_results(solver.h, os.fsencode(folder), entries, n, src_index))


def write_snapshots(solver, folder):
    """Final-state Snapshot_*.txt files of every field (tw_write_snapshots)."""
    solver._chk(lib().tw_write_snapshots(solver.h, os.fsencode(folder)))


def run_case(input_file, srcrcv_file, result_root, options=None, precision="fp32", cpst="0", model_dir=None,
             reference_boundaries=True, sources=None):
    """Run every source of a SrcRcv list as the reference's host driver does
    and write its result files.  reference_boundaries: run free-surface x /
    periodic y, as the reference's elastic solver does whatever the file's
    letters (the default); otherwise use the letters as given.  model_dir:
    read rho.bin, vp.bin, vs.bin there (inverse.cpp:6-130) instead of the
    driver's constant model.  Returns {source index: folder}."""
    inp = read_input_file(input_file, options)
    cfg = input_to_config(inp, ELASTIC, precision, cpst)
    if reference_boundaries:
        cfg.bc = b"FFPP"
    entries, n = read_srcrcv_file(srcrcv_file)
    mx, my = inp.model_m[0], inp.model_m[1]
    if model_dir is None:
        rho, vp, vs = (np.full((mx + 1, my + 1), v) for v in (1.0, 2.0, 1.0))
    else:
        rho, vp, vs = (read_model_file(os.path.join(model_dir, name + ".bin"), mx, my, inp.bc)
                       for name in ("rho", "vp", "vs"))
    folder = result_folder(inp, cfg, result_root)
    os.makedirs(folder, exist_ok=True)
    s = Solver(cfg)
    s.set_model(rho, vp, vs, inp.num_dx_model, inp.den_dx_model)
    out = {}
    src_ids = [entries[i].src_index for i in range(n) if entries[i].kind == 0]
    for k in (sources if sources is not None else src_ids):
        s.reset()
        setup_shot(s, inp, entries, n, k)
        s.step(inp.nt)
        write_results(s, folder, entries, n, k)
        out[k] = folder
    s.close()
    return out
