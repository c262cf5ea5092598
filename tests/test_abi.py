"""The C ABI library loads and exports every symbol include/namsa_hip.h declares.
CPU only: no compute call is made; argument checks that fail before any HIP
call are exercised."""
import ctypes
import re
import subprocess

import pytest

import namsa_amd._lib as L


def test_library_loads_with_expected_abi():
    lib = L.load()
    assert lib.namsa_abi_version() == L.ABI_VERSION


def test_every_declared_symbol_is_exported():
    declared = L.declared_functions()
    assert len(declared) >= 12
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (namsa_\w+)", out))
    assert set(declared) <= exported, set(declared) - exported
    assert set(declared) == set(L.SIGNATURES)


def test_abi_version_macro_matches():
    text = L.HEADER_PATH.read_text()
    assert int(re.search(r"#define NAMSA_ABI_VERSION (\d+)", text).group(1)) == L.ABI_VERSION


def test_supported_sizes():
    lib = L.load()
    for n in (128, 192, 256, 384, 512, 768, 1024, 2048):
        assert lib.namsa_fft_supported(n) == 1
    for n in (0, 100, 500, 4096):
        assert lib.namsa_fft_supported(n) == 0


def test_argument_errors_report_without_gpu():
    lib = L.load()
    rc = lib.namsa_fft2d(None, 1, 512, 512, 0, None)
    assert rc == -1 and b"null data" in lib.namsa_last_error()
    rc = lib.namsa_fft2d(ctypes.c_void_p(16), 1, 500, 512, 0, None)
    assert rc == -2 and b"500" in lib.namsa_last_error()
    rc = lib.namsa_multislice(None, None, 0, 512, 512, 8.0, None, 0, None, 0, 0, 0, 0, 1.0, None, None,
                              None, 0, None, None, None)
    assert rc == -1
    rc = lib.namsa_slice_cols(ctypes.c_void_p(16), 1, 512, 512, ctypes.c_void_p(16), 300, 200, None)
    assert rc == -1 and b"band" in lib.namsa_last_error()
    with pytest.raises(L.NamsaError):
        L.call("namsa_norm2", None, 0, None, None)
    assert lib.namsa_wave_tiled(2048, 2048) == 1 and lib.namsa_wave_tiled(512, 512) == 0
    rc = lib.namsa_slice_rows_tiled(ctypes.c_void_p(16), 1, 512, 512, ctypes.c_void_p(16), 0, 512, None)
    assert rc == -2 and b"tiled" in lib.namsa_last_error()
    rc = lib.namsa_slice_cols_tiled(ctypes.c_void_p(16), 1, 512, 512, ctypes.c_void_p(16), 0, 512, None)
    assert rc == -2 and b"tiled" in lib.namsa_last_error()
    assert lib.namsa_tile_plane(ctypes.c_void_p(16), ctypes.c_void_p(16), 512, 512, None) == -1
    assert lib.namsa_tile_plane(ctypes.c_void_p(16), ctypes.c_void_p(32), 511, 512, None) == -1
    fake = ctypes.c_void_p(16)
    rc = lib.namsa_multislice(fake, fake, 1, 512, 512, 8.0, fake, 1, fake, 0, 512, 1, L.OUT_CBED | L.PM_TILED, 1.0,
                              fake, fake, None, 0, None, None, None)
    assert rc == -1 and b"NAMSA_PM_TILED" in lib.namsa_last_error()
    # sized for the narrowest column tile of any schedule (4 columns)
    assert lib.namsa_det_workspace_doubles(4, 512, 512, 2) == 4 * (512 // 4) * 2
    assert lib.namsa_det_workspace_doubles(4, 500, 512, 2) == 0


def test_schedule_registry_host_side():
    from namsa_amd import autotune

    for n in (128, 192, 256, 384, 512, 768, 1024, 2048):
        for pass_ in (0, 1):
            names = autotune.schedules(n, pass_)
            assert len(names) >= 2 and len(set(names)) == len(names)
            keep = autotune.selected(n, pass_)
            autotune.select(n, pass_, names[-1])
            assert autotune.selected(n, pass_) == len(names) - 1
            autotune.select(n, pass_, keep)
    lib = L.load()
    assert lib.namsa_sched_count(500, 0) == 0
    assert lib.namsa_sched_name(512, 0, 99) is None
    assert lib.namsa_sched_set(512, 1, 99) == -1


def test_precise_schedule_policy():
    """Thick runs swap each compact-twiddle schedule for the schedule holding
    every twiddle it derives from, or for the partner its entry names (the
    multi-image column schedules, and the 2048 row and column schedules whose
    exact partner is a different kernel); other schedules map to themselves."""
    from namsa_amd import autotune

    lib = L.load()
    assert lib.namsa_precise_slices(512) == 300 and lib.namsa_precise_slices(768) == 200
    assert lib.namsa_precise_slices(500) == -1
    named = {(768, 1, "e12.r34444.c8.pg.sb.w4.ct"): "e12.r34444.c8.db.pf",
             (768, 1, "e12.r34444.c8.sb.w4.i2.ct"): "e12.r34444.c8.db.pf",
             (512, 1, "e8.r888.c8.pg.sb.w6.ct"): "e8.r888.c8.lr.pg.sb.w6",
             (768, 1, "e12.r12.4.4.4.c8.sb.w4.i2.ct"): "e12.r12.4.4.4.c8.db.pf",
             (1024, 1, "e16.r16.8.8.c8.sb.w4.i4.ct"): "e16.r16.8.8.c8.lr.pg.sb.w4",
             (2048, 0, "e16.r16.16.8.tr.ct"): "e16.r16.16.8.tr.ls",
             (2048, 1, "e16.r16.16.8.c4.sb.w4.i3.ct"): "e16.r16.16.8.c4.lr.pg.sb.w4"}
    for n in (128, 192, 256, 384, 512, 768, 1024, 2048):
        for p in (0, 1):
            names = autotune.schedules(n, p)
            for v, name in enumerate(names):
                q = lib.namsa_sched_precise(n, p, v)
                if (n, p, name) in named:
                    expect = names.index(named[(n, p, name)])
                else:
                    expect = names.index(name[:-3]) if name.endswith(".ct") else v
                assert q == expect, (n, p, name, names[q])
                assert not names[q].endswith(".ct"), (n, p, name, names[q])
    assert lib.namsa_sched_precise(512, 0, 99) < 0
    assert autotune.driver_schedules(512, 150) == (autotune.selected(512, 0), autotune.selected(512, 1))
