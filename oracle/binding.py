"""ctypes binding of the oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product (vct_amd/).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
_lib = None

F16 = C.c_float * 16
F3 = C.c_float * 3


class Scene(C.Structure):
    _fields_ = [("ntri", C.c_int), ("nobj", C.c_int), ("pos", C.c_void_p), ("nrm", C.c_void_p),
                ("tan", C.c_void_p), ("albedo", C.c_void_p), ("obj", C.c_void_p), ("rough", C.c_void_p),
                ("model", C.c_void_p), ("nmat", C.c_void_p), ("vtan", C.c_void_p), ("uv", C.c_void_p),
                ("texid", C.c_void_p), ("tex", C.c_void_p), ("ntex", C.c_int)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("levels", C.c_int), ("texels", C.c_void_p)]


class VoxParams(C.Structure):
    _fields_ = [("mvp_x", C.c_void_p), ("mvp_y", C.c_void_p), ("mvp_z", C.c_void_p), ("axis_override", C.c_int),
                ("lighting", C.c_int), ("msaa", C.c_int), ("dim", C.c_int), ("warpmap", C.c_void_p),
                ("accum", C.c_int), ("sums", C.c_void_p)]


class WarpParams(C.Structure):
    _fields_ = [("high_resolution", C.c_float), ("low_resolution", C.c_float), ("axes", C.c_int * 3),
                ("linear", C.c_int), ("use_weights_texture", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("total_voxel_fragments", C.c_uint32), ("unique_voxels", C.c_uint32),
                ("max_fragments_per_voxel", C.c_uint32), ("lit_voxels", C.c_uint32)]


class Cone(C.Structure):
    _fields_ = [("steps", C.c_int), ("angle", C.c_float), ("bias", C.c_float), ("h0", C.c_float),
                ("lod_offset", C.c_float)]


class TraceParams(C.Structure):
    _fields_ = [("diffuse", Cone), ("specular", Cone), ("enable_reflections", C.c_int),
                ("angle_from_roughness", C.c_int), ("draw_occlusion", C.c_int), ("has_roughness", C.c_int),
                ("ambient_scale", C.c_float), ("reflect_scale", C.c_float), ("eye", F3), ("dim", C.c_int),
                ("nlevels", C.c_int), ("center", F3), ("vmin", F3), ("vmax", F3), ("anisotropic", C.c_int),
                ("warpmap", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("eye", F3), ("view", F16), ("projection", F16),
                ("cull_back", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp = C.c_void_p
        sig = {
            "or_mul4": (None, [vp, vp, vp]),
            "or_ortho": (None, [C.c_float] * 6 + [vp]),
            "or_lookat": (None, [vp, vp, vp, vp]),
            "or_perspective": (None, [C.c_float] * 4 + [vp]),
            "or_inverse4": (C.c_int, [vp, vp]),
            "or_normal_matrix": (None, [vp, vp]),
            "or_mat4_vec4": (None, [vp, vp, vp]),
            "or_voxel_axis_projections": (None, [vp] * 6),
            "or_float_to_half": (C.c_uint16, [C.c_float]),
            "or_half_to_float": (C.c_float, [C.c_uint16]),
            "or_float_to_fp8": (C.c_uint8, [C.c_float]),
            "or_fp8_to_float": (C.c_float, [C.c_uint8]),
            "or_pack_unorm4x8": (C.c_uint32, [vp]),
            "or_shadowmap": (None, [C.POINTER(Scene), vp, vp]),
            "or_voxelize": (None, [C.POINTER(Scene), vp, C.c_int, vp, C.POINTER(VoxParams), vp, vp,
                                   C.POINTER(Stats)]),
            "or_normalize_inject": (None, [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, C.c_uint32, C.c_float,
                                           C.c_int, vp, vp, vp, C.POINTER(Stats)]),
            "or_normalize_inject_temporal": (None, [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, C.c_uint32, C.c_float,
                                                    C.c_float, C.c_int, vp, vp, vp, vp, C.POINTER(Stats)]),
            "or_warp_occupancy": (None, [C.POINTER(Scene), C.POINTER(VoxParams), vp]),
            "or_tess_levels": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
            "or_tess_count": (C.c_int, [vp]),
            "or_tess_point": (None, [vp, C.c_int, vp]),
            "or_voxelize_tess": (None, [C.POINTER(Scene), C.c_int, vp, vp, vp, C.c_int, vp, vp]),
            "or_warpmap": (None, [vp, C.POINTER(WarpParams), vp]),
            "or_warp_sample": (None, [vp, vp, vp]),
            "or_fill_holes": (None, [vp, vp, C.c_int, C.c_int]),
            "or_mip": (None, [vp, C.c_int, C.c_int, C.c_int]),
            "or_mip_aniso": (None, [vp, vp, C.c_int, C.c_int, C.c_int]),
            "or_fp8_quantize_level": (None, [vp, C.c_int]),
            "or_mip_albedo": (None, [vp, C.c_int, C.c_int]),
            "or_normalize_voxels": (None, [vp, C.c_int, C.c_float, vp, vp, vp, C.POINTER(Stats)]),
            "or_gbuffer": (None, [C.POINTER(Scene), C.POINTER(Camera), vp]),
            "or_trace": (None, [vp, C.c_int, vp, vp, C.POINTER(TraceParams), vp, vp]),
            "or_trace64": (None, [vp, C.c_int, vp, vp, C.POINTER(TraceParams), vp, vp]),
            "or_trace_cone": (None, [vp, vp, C.c_int, C.c_int, vp, vp, vp, C.POINTER(Cone), vp]),
            "or_sample": (None, [vp, C.c_int, C.c_int, vp, C.c_float, vp]),
            "or_calc_shadow_factor": (C.c_float, [vp, vp, C.c_int]),
            "or_compose": (None, [vp, C.c_int, vp, vp, vp, C.c_int, vp, C.c_int, C.c_int, vp, vp]),
            "or_tex_sample": (None, [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, vp]),
            "or_tex_rho2_affine": (C.c_float, [vp, vp, C.c_float, C.c_float]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- math
def ortho(l, r, b, t, n, f):
    o = np.zeros(16, np.float32)
    lib().or_ortho(l, r, b, t, n, f, _p(o))
    return o


def lookat(eye, center, up):
    o = np.zeros(16, np.float32)
    e, c, u = (np.asarray(v, np.float32) for v in (eye, center, up))
    lib().or_lookat(_p(e), _p(c), _p(u), _p(o))
    return o


def perspective(fovy, aspect, n, f):
    o = np.zeros(16, np.float32)
    lib().or_perspective(fovy, aspect, n, f, _p(o))
    return o


def inverse4(m):
    m = np.ascontiguousarray(m, np.float32)
    o = np.zeros(16, np.float32)
    ok = lib().or_inverse4(_p(m), _p(o))
    return o if ok else None


def mul4(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    o = np.zeros(16, np.float32)
    lib().or_mul4(_p(a), _p(b), _p(o))
    return o


def normal_matrices(model: np.ndarray) -> np.ndarray:
    model = np.ascontiguousarray(model, np.float32).reshape(-1, 16)
    out = np.zeros((model.shape[0], 9), np.float32)
    for i in range(model.shape[0]):
        lib().or_normal_matrix(_p(model[i]), _p(out[i]))
    return out


def voxel_projections(center, vmin, vmax):
    c, a, b = (np.asarray(v, np.float32) for v in (center, vmin, vmax))
    x, y, z = (np.zeros(16, np.float32) for _ in range(3))
    lib().or_voxel_axis_projections(_p(c), _p(a), _p(b), _p(x), _p(y), _p(z))
    return x, y, z


def half_round(x: np.ndarray) -> np.ndarray:
    f = np.vectorize(lambda v: lib().or_half_to_float(lib().or_float_to_half(float(v))), otypes=[np.float32])
    return f(np.asarray(x, np.float32))


def fp8_encode(x: float) -> int:
    return int(lib().or_float_to_fp8(float(x)))


def fp8_decode(c: int) -> float:
    return float(lib().or_fp8_to_float(int(c)))


# ---------------------------------------------------------------- scene
class SceneArrays:
    """Keeps the numpy arrays alive behind an or_scene."""

    def __init__(self, mesh: dict, model: np.ndarray):
        self.pos = np.ascontiguousarray(mesh["pos"], np.float32)
        self.nrm = np.ascontiguousarray(mesh["nrm"], np.float32)
        self.tan = np.ascontiguousarray(mesh["tan"], np.float32)
        self.albedo = np.ascontiguousarray(mesh["albedo"], np.uint32)
        self.obj = np.ascontiguousarray(mesh["obj"], np.uint32)
        self.rough = None if mesh.get("rough") is None else np.ascontiguousarray(mesh["rough"], np.float32)
        self.model = np.ascontiguousarray(model, np.float32).reshape(-1, 16)
        self.nmat = normal_matrices(self.model)
        self.vtan = None if mesh.get("vtan") is None else np.ascontiguousarray(mesh["vtan"], np.float32)
        # diffuse maps: mesh["uv"] [ntri][3][2], mesh["texid"] [ntri], mesh["textures"] a list of
        # (width, height, levels, texels uint32 array of every level)
        self.uv = None if mesh.get("uv") is None else np.ascontiguousarray(mesh["uv"], np.float32)
        self.texid = None if mesh.get("texid") is None else np.ascontiguousarray(mesh["texid"], np.int32)
        texs = mesh.get("textures") or []
        self.texels = [np.ascontiguousarray(t[3], np.uint32) for t in texs]
        self.tex = (Texture * max(1, len(texs)))(*[Texture(int(t[0]), int(t[1]), int(t[2]), _p(a))
                                                  for t, a in zip(texs, self.texels)])
        self.s = Scene(int(self.albedo.size), int(self.model.shape[0]), _p(self.pos), _p(self.nrm), _p(self.tan),
                       _p(self.albedo), _p(self.obj), None if self.rough is None else _p(self.rough),
                       _p(self.model), _p(self.nmat), None if self.vtan is None else _p(self.vtan),
                       None if self.uv is None else _p(self.uv), None if self.texid is None else _p(self.texid),
                       C.cast(self.tex, C.c_void_p), len(texs))


def tex_sample(textures, tid, s, t, rho2):
    """or_tex_sample: textures as in SceneArrays (mesh["textures"]); returns 4 floats."""
    texels = [np.ascontiguousarray(x[3], np.uint32) for x in textures]
    arr = (Texture * max(1, len(textures)))(*[Texture(int(x[0]), int(x[1]), int(x[2]), _p(a))
                                              for x, a in zip(textures, texels)])
    out = (C.c_float * 4)()
    lib().or_tex_sample(C.cast(arr, C.c_void_p), len(textures), int(tid), s, t, rho2, out)
    return np.array(out[:], np.float32)


def tex_rho2_affine(win, st, W, H) -> float:
    w = np.ascontiguousarray(win, np.float32).reshape(3, 2)
    a = np.ascontiguousarray(st, np.float32).reshape(6)
    return float(lib().or_tex_rho2_affine(_p(w), _p(a), W, H))


def _light_array(lights):
    from vct_amd import _abi as abi  # identical layout to or_light
    return (abi.Light * len(lights))(*lights)


def shadowmap(sc: SceneArrays, light) -> np.ndarray:
    from vct_amd import _abi as abi
    n = int(light.shadow_size)
    out = np.empty((n, n), np.float32)
    L = abi.Light.from_buffer_copy(light)
    lib().or_shadowmap(C.byref(sc.s), C.byref(L), _p(out))
    return out


def voxelize(sc: SceneArrays, lights, shadowmaps, vox, dim: int, msaa=None, lighting=None, axis_override=None,
             warpmap=None, accum=1, sums=None):
    """accum: 1 atomicMax (default), 0 the RGBA8 running average, 2 fixed-point
    sums into `sums` (uint32 [dim^3][8], linear x-fastest, zeroed here)."""
    mx = np.array(list(vox.mvp_x), np.float32)
    my = np.array(list(vox.mvp_y), np.float32)
    mz = np.array(list(vox.mvp_z), np.float32)
    wm = None if warpmap is None else np.ascontiguousarray(warpmap, np.float32)
    if accum == 2:
        assert sums is not None and sums.dtype == np.uint32 and sums.size == dim ** 3 * 8 and sums.flags.c_contiguous
        sums[...] = 0
    p = VoxParams(_p(mx), _p(my), _p(mz), vox.axis_override if axis_override is None else axis_override,
                  vox.lighting if lighting is None else lighting, vox.msaa if msaa is None else msaa, dim,
                  None if wm is None else _p(wm), int(accum), None if sums is None else _p(sums))
    la = _light_array(lights)
    sms = [np.ascontiguousarray(s, np.float32) if s is not None else None for s in shadowmaps]
    ptrs = (C.c_void_p * max(1, len(sms)))(*[(s.ctypes.data if s is not None else None) for s in sms])
    alb = np.zeros(dim ** 3, np.uint32)
    nrm = np.zeros(dim ** 3, np.uint32)
    st = Stats()
    lib().or_voxelize(C.byref(sc.s), C.cast(la, C.c_void_p), len(lights), C.cast(ptrs, C.c_void_p), C.byref(p),
                      _p(alb), _p(nrm), C.byref(st))
    return alb.reshape(dim, dim, dim), nrm.reshape(dim, dim, dim), st


def normalize_inject(albedo: np.ndarray, dim, center, vmin, vmax, lights, shadowmaps, light_mask=0xffffffff,
                     opacity=0.5, fp32=False, prev=None, decay=0.8, warpmap=None):
    """prev (float4 per voxel, last frame's stored level 0) selects the
    temporal filter with the given decay."""
    alb = np.ascontiguousarray(albedo, np.uint32).copy().reshape(-1)
    c, a, b = (np.asarray(v, np.float32) for v in (center, vmin, vmax))
    la = _light_array(lights)
    sms = [np.ascontiguousarray(s, np.float32) if s is not None else None for s in shadowmaps]
    ptrs = (C.c_void_p * max(1, len(sms)))(*[(s.ctypes.data if s is not None else None) for s in sms])
    rad = np.zeros((dim ** 3, 4), np.float32)
    lm = np.zeros(dim ** 3, np.uint8)
    st = Stats()
    wm = None if warpmap is None else np.ascontiguousarray(warpmap, np.float32)
    wp = None if wm is None else _p(wm)
    if prev is None:
        lib().or_normalize_inject(_p(alb), dim, _p(c), _p(a), _p(b), C.cast(la, C.c_void_p), len(lights),
                                  C.cast(ptrs, C.c_void_p), light_mask, opacity, 1 if fp32 else 0, wp, _p(rad),
                                  _p(lm), C.byref(st))
    else:
        pv = np.ascontiguousarray(prev, np.float32).reshape(-1, 4)
        assert pv.shape[0] == dim ** 3
        lib().or_normalize_inject_temporal(_p(alb), dim, _p(c), _p(a), _p(b), C.cast(la, C.c_void_p), len(lights),
                                           C.cast(ptrs, C.c_void_p), light_mask, opacity, decay,
                                           1 if fp32 else 0, wp, _p(pv), _p(rad), _p(lm), C.byref(st))
    return alb.reshape(dim, dim, dim), rad.reshape(dim, dim, dim, 4), lm, st


def normalize_voxels(sums: np.ndarray, dim: int, opacity: float = 0.5):
    """normalizeVoxels.comp on fixed-point sums: (albedo, normal, rad0, stats)."""
    alb = np.zeros(dim ** 3, np.uint32)
    nrm = np.zeros(dim ** 3, np.uint32)
    rad = np.zeros((dim ** 3, 4), np.float32)
    st = Stats()
    lib().or_normalize_voxels(_p(np.ascontiguousarray(sums, np.uint32)), dim, opacity, _p(alb), _p(nrm), _p(rad),
                              C.byref(st))
    return alb.reshape(dim, dim, dim), nrm.reshape(dim, dim, dim), rad.reshape(dim, dim, dim, 4), st


def fill_holes(rad0: np.ndarray, fp32=False) -> np.ndarray:
    src = np.ascontiguousarray(rad0, np.float32)
    dim = src.shape[0]
    dst = np.empty_like(src)
    lib().or_fill_holes(_p(src), _p(dst), dim, 1 if fp32 else 0)
    return dst


def mip(rad0: np.ndarray, nlevels: int, fp32=False) -> list:
    dim = rad0.shape[0]
    levels = [np.ascontiguousarray(rad0, np.float32)]
    for l in range(1, nlevels):
        s = max(1, dim >> l)
        levels.append(np.zeros((s, s, s, 4), np.float32))
    ptrs = (C.c_void_p * nlevels)(*[a.ctypes.data for a in levels])
    lib().or_mip(C.cast(ptrs, C.c_void_p), dim, nlevels, 1 if fp32 else 0)
    return levels


def mip_aniso(rad0: np.ndarray, nlevels: int, fp32=False) -> list:
    """Returns [None, faces_l1, faces_l2, ...] where faces_l = list of 6 volumes."""
    dim = rad0.shape[0]
    src = np.ascontiguousarray(rad0, np.float32)
    faces = [None]
    rows = [None]
    for l in range(1, nlevels):
        s = max(1, dim >> l)
        fl = [np.zeros((s, s, s, 4), np.float32) for _ in range(6)]
        faces.append(fl)
        rows.append((C.c_void_p * 6)(*[a.ctypes.data for a in fl]))
    top = (C.c_void_p * nlevels)(*([None] + [C.cast(r, C.c_void_p).value for r in rows[1:]]))
    lib().or_mip_aniso(_p(src), C.cast(top, C.c_void_p), dim, nlevels, 1 if fp32 else 0)
    return faces


def mip_albedo(grid: np.ndarray, nlevels: int) -> list:
    """The albedo (voxelColor) chain: [grid, level 1, ...] as u32 RGBA8 volumes."""
    dim = grid.shape[0]
    levels = [np.ascontiguousarray(grid, np.uint32).reshape(dim, dim, dim)]
    for l in range(1, nlevels):
        s = max(1, dim >> l)
        levels.append(np.zeros((s, s, s), np.uint32))
    ptrs = (C.c_void_p * nlevels)(*[a.ctypes.data for a in levels])
    lib().or_mip_albedo(C.cast(ptrs, C.c_void_p), dim, nlevels)
    return levels


def fp8_quantize(level: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(level, np.float32).copy()
    lib().or_fp8_quantize_level(_p(a), a.shape[0])
    return a


def gbuffer(sc: SceneArrays, cam) -> np.ndarray:
    c = Camera.from_buffer_copy(cam)
    out = np.zeros((cam.height, cam.width, 16), np.float32)
    lib().or_gbuffer(C.byref(sc.s), C.byref(c), _p(out))
    return out


def trace_params(tr, dim, nlevels, center, vmin, vmax, has_roughness, anisotropic=0, warpmap=None) -> TraceParams:
    """warpmap: a float32 [32^3 x 4] array the caller keeps alive while tracing."""
    d, s = tr.diffuse, tr.specular
    return TraceParams(Cone(d.steps, d.cone_angle, d.bias, d.cone_initial_height, d.lod_offset),
                       Cone(s.steps, s.cone_angle, s.bias, s.cone_initial_height, s.lod_offset),
                       tr.enable_reflections, tr.cone_angle_from_roughness, tr.draw_occlusion,
                       1 if has_roughness else 0, tr.ambient_scale, tr.reflect_scale, F3(*tr.eye), dim, nlevels,
                       F3(*center), F3(*vmin), F3(*vmax), anisotropic, None if warpmap is None else _p(warpmap))


def _pyramid_ptrs(levels, aniso_faces):
    lp = (C.c_void_p * len(levels))(*[np.ascontiguousarray(a).ctypes.data for a in levels])
    if aniso_faces is None:
        return lp, None, None
    rows = [None] + [(C.c_void_p * 6)(*[f.ctypes.data for f in aniso_faces[l]]) for l in range(1, len(levels))]
    top = (C.c_void_p * len(levels))(*([None] + [C.cast(r, C.c_void_p).value for r in rows[1:]]))
    return lp, top, rows


def trace(gbuf: np.ndarray, levels, params: TraceParams, aniso_faces=None, fp64=False):
    g = np.ascontiguousarray(gbuf, np.float32)
    npix = g.size // 16
    lp, top, _keep = _pyramid_ptrs(levels, aniso_faces)
    dt = np.float64 if fp64 else np.float32
    ind = np.zeros((npix, 4), dt)
    spec = np.zeros((npix, 4), dt)
    fn = lib().or_trace64 if fp64 else lib().or_trace
    fn(_p(g), npix, C.cast(lp, C.c_void_p), None if top is None else C.cast(top, C.c_void_p), C.byref(params),
       _p(ind), _p(spec))
    return ind, spec


def shadow_factor(lsp, shadowmap: np.ndarray) -> float:
    lp = np.ascontiguousarray(lsp, np.float32)
    sm = np.ascontiguousarray(shadowmap, np.float32)
    return float(lib().or_calc_shadow_factor(_p(lp), _p(sm), sm.shape[0]))


def compose(gbuf: np.ndarray, indirect, specular, lights, shadowmaps, params, has_roughness=False,
            occluded=True) -> np.ndarray:
    """params: a vct_amd._abi.ComposeParams (same layout as or_compose_params).
    indirect/specular: float4 per pixel (or None when not used)."""
    from vct_amd import _abi as abi
    g = np.ascontiguousarray(gbuf, np.float32)
    npix = g.size // 16
    ind = np.zeros((npix, 4), np.float32) if indirect is None else np.ascontiguousarray(indirect, np.float32)
    spe = np.zeros((npix, 4), np.float32) if specular is None else np.ascontiguousarray(specular, np.float32)
    assert ind.size == npix * 4 and spe.size == npix * 4
    la = _light_array(lights) if lights else None
    sms = [None if s is None else np.ascontiguousarray(s, np.float32) for s in (shadowmaps or [])]
    sms += [None] * (len(lights) - len(sms))
    ptrs = (C.c_void_p * max(1, len(sms)))(*[(s.ctypes.data if s is not None else None) for s in sms])
    P = abi.ComposeParams.from_buffer_copy(params)
    out = np.zeros((npix, 4), np.float32)
    lib().or_compose(_p(g), npix, _p(ind), _p(spe), None if la is None else C.cast(la, C.c_void_p), len(lights),
                     C.cast(ptrs, C.c_void_p), 1 if has_roughness else 0, 1 if occluded else 0, C.byref(P), _p(out))
    return out


# ---------------------------------------------------------------- tessellation voxelizer
def tess_levels(P, dim, center, vmin, vmax):
    """(point count, [inner, outer0, outer1, outer2]) of one patch (world positions P[3][3])."""
    Pa = np.ascontiguousarray(P, np.float32).reshape(3, 3)
    c, mn, mx = (np.ascontiguousarray(v, np.float32) for v in (center, vmin, vmax))
    lv = np.zeros(4, np.int32)
    n = lib().or_tess_levels(_p(Pa), int(dim), _p(c), _p(mn), _p(mx), _p(lv))
    return int(n), [int(x) for x in lv]


def tess_points(lv) -> np.ndarray:
    """Barycentric (u, v, w) of every point of a patch with levels lv, [count][3]."""
    lv = np.ascontiguousarray(lv, np.int32)
    n = lib().or_tess_count(_p(lv))
    out = np.zeros((n, 3), np.float32)
    for j in range(n):
        lib().or_tess_point(_p(lv), j, out[j].ctypes.data)
    return out


def voxelize_tess(sc: SceneArrays, dim, center, vmin, vmax, accum=1):
    """Albedo and normal grids (z, y, x) of the tessellation voxelizer."""
    c, mn, mx = (np.ascontiguousarray(v, np.float32) for v in (center, vmin, vmax))
    alb = np.zeros(dim ** 3, np.uint32)
    nrm = np.zeros(dim ** 3, np.uint32)
    lib().or_voxelize_tess(C.byref(sc.s), int(dim), _p(c), _p(mn), _p(mx), int(accum), _p(alb), _p(nrm))
    return alb.reshape(dim, dim, dim), nrm.reshape(dim, dim, dim)


# ---------------------------------------------------------------- warp texture
def warp_params(high=2.0, low=0.5, axes=(1, 1, 1), linear=0, use_weights_texture=1) -> WarpParams:
    return WarpParams(high, low, (C.c_int * 3)(*axes), linear, use_weights_texture)


def warp_occupancy(sc: SceneArrays, vox, msaa=None) -> np.ndarray:
    """32^3 occupancy (z, y, x) of the scene, voxelize.frag:187-193."""
    mx, my, mz = (np.array(list(getattr(vox, m)), np.float32) for m in ("mvp_x", "mvp_y", "mvp_z"))
    p = VoxParams(_p(mx), _p(my), _p(mz), vox.axis_override, 0, vox.msaa if msaa is None else msaa, 32, None)
    occ = np.zeros(32 ** 3, np.uint32)
    lib().or_warp_occupancy(C.byref(sc.s), C.byref(p), _p(occ))
    return occ.reshape(32, 32, 32)


def warpmap(occ: np.ndarray, params: WarpParams | None = None) -> np.ndarray:
    o = np.ascontiguousarray(occ, np.uint32).reshape(-1)
    assert o.size == 32 ** 3
    out = np.zeros((32, 32, 32, 4), np.float32)
    lib().or_warpmap(_p(o), C.byref(params or warp_params()), _p(out))
    return out


def warp_sample(wm: np.ndarray, tc) -> np.ndarray:
    w = np.ascontiguousarray(wm, np.float32)
    t = np.asarray(tc, np.float32).reshape(3)
    out = np.zeros(3, np.float32)
    lib().or_warp_sample(_p(w), _p(t), _p(out))
    return out
