"""The C ABI without a GPU: exported symbols, struct layouts, argument errors.

Checks that libvct.so / libvct_app.so export every function include/*.h
declares, that the ctypes mirror matches the C struct layouts (a small C
program compiled against the headers prints the sizes and offsets), and that
invalid arguments are rejected before any HIP call."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

import vct_amd
from vct_amd import _abi as abi

ROOT = Path(__file__).resolve().parents[1]


def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vct_\w+)\s*\(", text)) - {"vct_ctx", "vct_app"})


@pytest.mark.parametrize("header,lib", [("vct.h", "libvct.so"), ("vct_app.h", "libvct_app.so")])
def test_library_exports_every_declared_function(header, lib):
    names = _declared(ROOT / "include" / header)
    assert len(names) > 10
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "vct_amd" / "_lib" / lib)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\b[TW] (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_ctypes_mirror_covers_every_function():
    names = set(_declared(ROOT / "include" / "vct.h"))
    assert names == {n for n, _, _ in abi.VCT_FUNCS}
    names = set(_declared(ROOT / "include" / "vct_app.h"))
    assert names == {n for n, _, _ in abi.APP_FUNCS}


def test_struct_layouts_match_headers(tmp_path):
    structs = {"vct_grid_desc": abi.GridDesc, "vct_mesh": abi.Mesh, "vct_light": abi.Light,
               "vct_voxelize_params": abi.VoxelizeParams, "vct_inject_params": abi.InjectParams,
               "vct_mip_params": abi.MipParams, "vct_camera": abi.Camera, "vct_cone_settings": abi.ConeSettings,
               "vct_trace_params": abi.TraceParams, "vct_stats": abi.Stats, "vct_pass_times": abi.PassTimes,
               "vct_app_config": abi.AppConfig, "vct_learned_params": abi.LearnedParams,
               "vct_compose_params": abi.ComposeParams, "vct_warpmap_params": abi.WarpmapParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vct_app.h"', "int main(void){"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append('printf("vct_gbuffer_texel %zu\\n", sizeof(vct_gbuffer_texel)); return 0;}')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                          check=True).stdout.split("\n") if l)
    for name, cls in structs.items():
        assert int(got[name]) == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, f"{name}.{f}"
    assert int(got["vct_gbuffer_texel"]) == abi.GBUFFER_TEXEL_BYTES


def _desc(**kw):
    d = dict(dim=64, levels=7, center=(0, 0, 0), vmin=(-1, -1, -1), vmax=(1, 1, 1), storage=0, anisotropic=0,
             device=0)
    d.update(kw)
    return abi.GridDesc(d["dim"], d["levels"], (C.c_float * 3)(*d["center"]), (C.c_float * 3)(*d["vmin"]),
                        (C.c_float * 3)(*d["vmax"]), d["storage"], d["anisotropic"], d["device"])


@pytest.mark.parametrize("kw,needle", [(dict(dim=100), "power of two"), (dict(dim=4), "power of two"),
                                       (dict(dim=1024), "power of two"), (dict(storage=7), "storage"),
                                       (dict(vmin=(1, -1, -1), vmax=(-1, 1, 1)), "extent")])
def test_create_rejects_invalid_descriptions_without_a_gpu(kw, needle):
    lib = vct_amd.load_library()
    h = C.c_void_p()
    rc = lib.vct_create(C.byref(h), C.byref(_desc(**kw)))
    assert rc == abi.VCT_ERR_INVALID
    assert needle in lib.vct_last_error(None).decode()
    assert not h.value


def test_null_arguments_are_rejected():
    lib = vct_amd.load_library()
    assert lib.vct_create(None, C.byref(_desc())) == abi.VCT_ERR_INVALID
    assert lib.vct_create(C.byref(C.c_void_p()), None) == abi.VCT_ERR_INVALID
    for fn in ("vct_shadowmap", "vct_fill_holes", "vct_sync"):
        args = [None, 0] if fn == "vct_shadowmap" else [None]
        assert getattr(lib, fn)(*args) == abi.VCT_ERR_INVALID
    assert lib.vct_voxelize(None, None) == abi.VCT_ERR_INVALID
    assert lib.vct_buffer_bytes(None, abi.VCT_BUF_ALBEDO, 0) == 0
    lib.vct_destroy(None)  # a no-op


def test_abi_version():
    assert vct_amd.load_library().vct_abi_version() == 1


def test_app_without_device_builds_scene_only():
    app = vct_amd.App(1, device=-1)
    assert app.ctx is None
    with pytest.raises(vct_amd.VCTError):
        app.render(0)
