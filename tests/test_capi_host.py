"""Host-only checks of the C ABI (no GPU needed): the library loads, exports
every function include/pic2d.h declares, and the parameters.h loader evaluates
the reference's macros as the reference's compiler would."""
import ctypes
import math
import os

import numpy as np
import pytest

import pic2d

REF = "/root/reference"


def test_library_exports_every_declared_symbol():
    lib = pic2d.load_library()
    names = pic2d.declared_functions()
    assert len(names) >= 35 and "pic_last_error" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_match_reference_macros():
    p = pic2d.default_params()
    q, me, amu, eps = 1.602177e-19, 9.10939e-31, 1.66054e-27, 8.854e-12
    assert p.nodes == 257 and p.init_particle_no == 512 * 512
    debye = math.sqrt(eps * 1.5 / (1e15 * q))
    assert p.length == 64 * debye
    wpe = math.sqrt(1e15 * 2.566971e-38 / (eps * me))
    assert p.dt == 0.1 / wpe and p.dt_dti == 1
    assert p.vte == math.sqrt(1.5 * q / me)
    assert p.ion_mass == 39.948 * amu
    assert p.neutral_density == (0.2 * 133.32) / (0.025 * q)
    assert p.run_time == 300 * int(1.0 / (2.0 * p.dt * 3.141592654 * 13.6e6)) + 5
    assert p.wall == pic2d.WALL_STEADY and p.mode == pic2d.MODE_REF


def test_loader_evaluates_c_expressions(tmp_path):
    f = tmp_path / "parameters.h"
    f.write_text("""
#define _CRT_SECURE_NO_WARNINGS
#ifndef parameters_h
#define parameters_h
#include "constants.h"
#define RAND_SEED 1234
#define X_NO_OF_CELLS 65 // 33
#define Y_NO_OF_CELLS X_NO_OF_CELLS
/* block
   comment */
#define TEMPERATURE  (double)2.0
#define VTE sqrt(TEMPERATURE*CHARGE/ELECTRON_MASS)
#define ION_MASS ((double)39.948*AMU)
#define XLENGTH ((double)0.01)
#define YLENGTH XLENGTH
#define INIT_PARTICLE_NO 640*640
#define DT ((double)0.1/sqrt((1.0e15*ELECTRON_CHARGE2)/(EPSILON*ELECTRON_MASS)))
#define DTI ((double)4.0*DT)
#define DT_DTI (int)(DTI/DT)
#define STEADYSTATE "NO"
#define HAVE_COLLISIONS false
#define SPECIES "HYDROGEN"
#define TOTAL_RUN_TIME ((int)7*(int)3 + 5)/2
#define PIC_MODE "PHYSICAL"
#define PIC_LOADER "RANDOM"
#define PIC_MG_SWEEPS 3
#endif
""")
    p = pic2d.load_params(f)
    assert p.nodes == 65 and p.seed == 1234
    assert p.vte == math.sqrt(2.0 * 1.602177e-19 / 9.10939e-31)
    assert p.init_particle_no == 409600 and p.dt_dti == 4
    assert p.run_time == 13              # integer division, as C
    assert p.wall == pic2d.WALL_ABSORB and p.have_collisions == 0
    assert p.gas == pic2d.GAS_HYDROGEN and p.mode == pic2d.MODE_PHYSICAL
    assert p.loader == pic2d.LOAD_RANDOM and p.length == 0.01
    assert p.mg_sweeps == 3


def test_loader_particle_count_control(tmp_path):
    """parameters.h:320-366 as pic.cpp:257-327 uses them (three groups)"""
    f = tmp_path / "parameters.h"
    f.write_text("""
#define TEMPERATURE  (double)2.0
#define VTE sqrt(TEMPERATURE*CHARGE/ELECTRON_MASS)
#define ION_MASS ((double)39.948*AMU)
#define VTI sqrt(0.05*CHARGE/ION_MASS)
#define STEADYSTATE "NO"
#define NUMBER_TO_INJECT 12
#define NO_GROUPS (int)3
#define UPPER_GROUP_LIMIT (int)4000
#define UPPER_GROUP_LIMIT2 (int)1000
#define NEW_GROUP_SIZE (int)2000
#define NEW_GROUP_SIZE2 (int)500
#define UPPER_V_LIMIT ((double)100.0*VTE)
#define VE1 (double)1.5*VTE
#define VI2 (double)4.0*VTI
#define SIZE_TO_CHECK (int)77
#define SIZE_TO_CHECK_E (int)99
""")
    p = pic2d.load_params(f)
    vte, vti = p.vte, p.vti
    assert p.reduce == 1 and p.number_to_inject == 12 and p.no_groups == 3
    assert list(p.size_to_check) == [99, 77]
    assert [p.v_limit[0][g] for g in range(3)] == [1.5 * vte, 3.0 * vte, 100.0 * vte]
    assert [p.v_limit[1][g] for g in range(3)] == [2.0 * vti, 4.0 * vti, 1000.0 * vti]
    assert [p.group_limit[0][g] for g in range(3)] == [4000] * 3
    assert [p.new_group_size[0][g] for g in range(3)] == [2000] * 3
    assert [p.group_limit[1][g] for g in range(3)] == [4000, 1000, 1000]
    assert [p.new_group_size[1][g] for g in range(3)] == [2000, 500, 500]
    # steady-state walls: the reference never reduces (pic.cpp:1132)
    f.write_text('#define STEADYSTATE "YES"\n')
    assert pic2d.load_params(f).reduce == 0
    f.write_text('#define STEADYSTATE "YES"\n#define PIC_REDUCE 1\n')
    assert pic2d.load_params(f).reduce == 1
    f.write_text('#define E_IMPACT_SEE "YES"\n#define WORK_FUNCTION (double)4.5\n')
    p = pic2d.load_params(f)
    assert p.e_impact_see == 1 and p.reflection == 0 and p.work_function == 4.5
    f.write_text('#define NO_GROUPS 4\n')
    with pytest.raises(pic2d.PicError) as e:
        pic2d.load_params(f)
    assert e.value.code == -7


def test_loader_errors(tmp_path):
    with pytest.raises(pic2d.PicError) as e:
        pic2d.load_params(tmp_path / "missing.h")
    assert e.value.code == -3
    f = tmp_path / "bad.h"
    f.write_text("#define X_NO_OF_CELLS (65 +\n")
    with pytest.raises(pic2d.PicError) as e:
        pic2d.load_params(f)
    assert e.value.code == -4
    f.write_text('#define X_NO_OF_CELLS 65\n#define Y_NO_OF_CELLS 33\n')
    with pytest.raises(pic2d.PicError) as e:
        pic2d.load_params(f)
    assert e.value.code == -7             # mglin needs a square grid


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
@pytest.mark.parametrize("rel", ["parameters.h", "parameter file for Ar steadystate/parameters.h",
                                 "parameter file for Ar with coll-boundary/parameters.h",
                                 "parameter file for H with coll-boundary/parameters.h"])
def test_reference_parameter_files_evaluate(rel):
    path = os.path.join(REF, rel)
    if not os.path.exists(path):
        pytest.skip("file absent")
    p = pic2d.load_params(path)
    d = pic2d.default_params()
    assert p.nodes == d.nodes and p.length == d.length and p.dt == d.dt
    if "coll-boundary" in rel:            # parameters.h:141-152 switched on
        assert p.e_impact_see and p.reflection and p.backscatter
        assert p.ion_impact_see == ("Ar " in rel)   # the H file leaves it "NO"
        assert p.wall == pic2d.WALL_ABSORB and p.work_function == 4.6


def test_create_reports_grid_and_device_errors():
    p = pic2d.default_params()
    p.nodes = 100
    with pytest.raises(pic2d.PicError) as e:
        pic2d.Pic(p)
    assert e.value.code == -2          # "n-1 must be a power of 2" (mglin.cpp:25-27)
    p = pic2d.default_params()
    assert p.mg_sweeps == -1           # chosen by the library
    p.mg_sweeps = 4
    with pytest.raises(pic2d.PicError) as e:
        pic2d.Pic(p)
    assert e.value.code == -1 and "mg_sweeps" in str(e.value)
    with pytest.raises(pic2d.PicError) as e:
        pic2d.mglin(np.zeros((6, 6)))
    assert e.value.code == -2


def test_estimate_and_serve_arguments():
    import workloads
    small, big = workloads.plain(), workloads.argon_256()
    es, eb = pic2d.estimate_bytes(small), pic2d.estimate_bytes(big)
    assert 0 < es < eb
    # 13.1 M particles per species, double-buffered 28 B each plus staging: > 1.4 GB
    assert eb > 2 * 2 * 13107200 * 28
    res, st = pic2d.serve([], int(1e9))
    assert res == [] and st["particle_steps"] == 0
    with pytest.raises(pic2d.PicError) as e:
        pic2d.serve([(small, 1, 0.0)], 0)
    assert e.value.code == -1
    with pytest.raises(pic2d.PicError) as e:       # can never be admitted: refused up front
        pic2d.serve([(big, 1, 0.0)], eb // 2)
    assert e.value.code == -6


def test_mglin_noalloc_rejects_bad_sizes_before_any_device_call():
    """the no-allocation FP64 multigrid checks its arguments on the host: a
    grid beyond its static scratch array (n > 1025) is PIC_ERR_CAPACITY, a
    grid that is not 2^k + 1 PIC_ERR_GRID (no GPU needed for either)"""
    lib = pic2d.load_library()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: the checks come first
    assert lib.pic_mglin_noalloc(fake, 2049, 2, 1.0, None) == -6
    assert b"1025" in lib.pic_last_error(None)
    assert lib.pic_mglin_noalloc(fake, 1000, 2, 1.0, None) == -2
    assert lib.pic_mglin_noalloc(None, 65, 2, 1.0, None) == -1
    assert lib.pic_mglin_noalloc(fake, 65, 0, 1.0, None) == -1


def test_server_api_host_logic():
    """the serving object without a GPU: argument checks, a request that can
    never fit the budget, unknown ids, cancel of a queued request before any
    device work (max_active 1 with the slot held by a request that cannot
    start here)"""
    lib = pic2d.load_library()
    s = ctypes.c_void_p()
    assert lib.pic_server_create(0, 4, ctypes.byref(s)) == -1
    assert lib.pic_server_create(1 << 30, 0, ctypes.byref(s)) == -1
    p = pic2d.default_params()
    with pic2d.Server(1 << 20, 1) as srv:        # 1 MiB: nothing fits
        with pytest.raises(pic2d.PicError) as e:
            srv.submit(p, 10)
        assert e.value.code == -6
        with pytest.raises(pic2d.PicError):
            srv.poll(12345)
    with pic2d.Server(1 << 40, 1) as srv:
        a = srv.submit(p, 10)
        st = srv.wait(a, timeout=120)
        # without a GPU the run fails at creation; with one it completes
        assert st["state"] in (pic2d.REQ_DONE, pic2d.REQ_FAILED)
        b = srv.submit(p, 10)
        srv.cancel(b)
        st = srv.wait(b, timeout=120)
        assert st["state"] in (pic2d.REQ_CANCELLED, pic2d.REQ_DONE, pic2d.REQ_FAILED)


def test_checkpoint_params_rejects_other_files(tmp_path):
    """pic_checkpoint_params reads only the header (no device call): a text
    file, an empty file and a truncated header are PIC_ERR_FORMAT, a missing
    file PIC_ERR_FILE"""
    bad = tmp_path / "x.ckpt"
    for data in (b"", b"0.1 0.2 0.3 0.4 0.5 1.0\n" * 10, b"PIC2DCK1" + b"\0" * 16):
        bad.write_bytes(data)
        with pytest.raises(pic2d.PicError) as e:
            pic2d.checkpoint_params(bad)
        assert e.value.code == pic2d.PIC_ERR_FORMAT
    with pytest.raises(pic2d.PicError) as e:
        pic2d.checkpoint_params(tmp_path / "missing.ckpt")
    assert e.value.code == pic2d.PIC_ERR_FILE
