"""The C-ABI library: it loads, exports every entry point include/photons_fmm.h declares,
and refuses to run without a GPU instead of computing anything on the CPU."""
import ctypes as C
import os
import subprocess

import pytest

import photons_gpu as pg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(pg.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "photons-2.0_amd")], check=True)
    return pg.lib()


def test_every_declared_symbol_is_exported(lib):
    names = pg.declared_symbols()
    # the drop-in layer and the native API are both there
    for must in ("fmm_prepare", "fmm_task", "fmm_ext", "p2m", "m2l", "l2p", "p2p_kernel", "walk_task_p2p_ext",
                 "photons_create", "photons_force_device", "photons_p2m_batch"):
        assert must in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_globals_are_exported(lib):
    for name in ("NPART", "MAXLEAF", "first_leaf", "first_node", "last_leaf", "last_node"):
        C.c_int.in_dll(lib, name)
    for name in ("MASSPART", "SoftenScale", "splitRadius", "cutoffRadius", "open_angle", "BOXSIZE"):
        C.c_double.in_dll(lib, name)
    for name in ("part", "leaf", "btree"):
        C.c_void_p.in_dll(lib, name)


def test_parameters_are_validated(lib):
    import ic
    _, prm = ic.workload_lattice32()
    bad = dict(prm.as_dict())
    bad["order"] = 7
    h = C.c_void_p()
    p = pg.make_params(bad)
    assert lib.photons_create(C.byref(h), C.byref(p), 0, None) != 0
    assert b"order" in lib.photons_last_error()


def test_no_gpu_means_an_error_not_a_fallback(lib):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    import ic
    _, prm = ic.workload_lattice32()
    with pytest.raises(pg.PhotonsError):
        pg.Context(prm)


def test_kick_drift_factors_match_the_oracle():
    # host-side arithmetic of the library (no GPU needed)
    import numpy as np
    import oracle
    for ai, af in ((0.02, 0.025), (0.3, 0.31), (0.5, 1.0)):
        li, lf = np.log(ai), np.log(af)
        assert abs(pg.kick_factor(li, lf) / oracle.kick_loga(li, lf) - 1) < 1e-14
        assert abs(pg.drift_factor(li, lf) / oracle.drift_loga(li, lf) - 1) < 1e-14


def test_c_driver_fails_loudly_without_a_gpu(tmp_path):
    """the compiled reference-style driver links the library; with no GPU the force
    call stops with a message instead of computing anything on the CPU"""
    import numpy as np
    exe = os.path.join(ROOT, "photons-2.0_amd", "examples", "fmm_driver")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array([100, 32, 0, 0], np.int32).tobytes())
        f.write(np.array([1000.0, 1.0, 1.0, 10.0, 60.0, 0.4]).tobytes())
        f.write(np.random.default_rng(0).random((100, 3)).tobytes())
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "GPU" in (r.stdout + r.stderr)
    # the message stays readable through photons_gpu_last_error() after the exit
    assert "fmm_driver: the force path stopped:" in r.stderr and "GPU" in r.stderr.split("stopped:")[1]


def test_run_program_fails_loudly_without_a_gpu(tmp_path):
    """the reference-style run program reads the parameter file and the IC, then stops
    with a message when there is no GPU; a bad parameter file stops it earlier"""
    import numpy as np
    import photons_gpu as pg
    exe = os.path.join(ROOT, "photons-2.0_amd", "examples", "photons_run")
    if not os.path.exists(exe):
        pytest.skip("run program not built")
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    pos = np.random.default_rng(0).random((4096, 3)) * 1000.0
    ic_path = tmp_path / "ic.gdt2"
    pg.snapshot_write(ic_path, pos, np.zeros_like(pos), boxsize=1000.0, redshift=9.0, mass=1.0)
    prm = tmp_path / "run.param"
    prm.write_text(f"InputPath {ic_path}\nOutputPath {tmp_path}\nOutputName out\nNumMeshSide 16\n")
    r = subprocess.run([exe, str(prm), "--nstep", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "GPU" in (r.stdout + r.stderr)
    assert not (tmp_path / "out_3.0").exists()
    prm.write_text(f"InputPath {ic_path}\nSnapFormat -2\nNumMeshSide 16\n")
    r = subprocess.run([exe, str(prm)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "SnapFormat" in r.stderr


def test_pm_mesh_index_bound(lib):
    """the PM's CIC corner keys are 32-bit mesh indices: a mesh with N^3 >= 2^32 cells
    (N >= 1626) is refused before anything touches the GPU"""
    h = C.c_void_p()
    lib.photons_pm_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_int, C.c_void_p]
    lib.photons_pm_last_error.restype = C.c_char_p
    for nside in (1626, 1700, 2048):
        assert lib.photons_pm_create(C.byref(h), nside, 1000.0, 10.0, 1.0, 0, None) != 0
        assert b"1624" in lib.photons_pm_last_error()
