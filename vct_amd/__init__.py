"""vct_amd -- MI355X-native voxel cone tracing.

The hot path of the reference renderer (sfreed141/vct): dominant-axis
voxelization, shadow-map radiance injection, the radiance mip chain and the
cone-traced indirect light, as hand-written gfx950 HIP kernels behind the C ABI
of include/vct.h (libvct.so), driven by the headless application layer of
include/vct_app.h (libvct_app.so).

This module is a thin ctypes binding for tests and benchmarks.  It performs no
computation itself: every pass runs in libvct.so on the GPU, and loading fails
loudly when the built library is missing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _abi as abi

__all__ = ["VCTError", "load_library", "load_app_library", "Context", "App", "abi"]

LIB_DIR = Path(__file__).resolve().parent / "_lib"
_lib = None
_app = None


class VCTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load_library() -> C.CDLL:
    """dlopen libvct.so (built by vct_amd/build.py). Raises if it is missing."""
    global _lib
    if _lib is None:
        path = LIB_DIR / "libvct.so"
        if not path.exists():
            raise ImportError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
        abi.bind(lib, abi.VCT_FUNCS)
        _lib = lib
    return _lib


def load_app_library() -> C.CDLL:
    global _app
    if _app is None:
        load_library()
        path = LIB_DIR / "libvct_app.so"
        if not path.exists():
            raise ImportError(f"{path} is missing: build it first")
        lib = C.CDLL(str(path))
        abi.bind(lib, abi.APP_FUNCS)
        _app = lib
    return _app


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Context:
    """Owns one vct_ctx (the VCT resources of src/Application.h:107-156)."""

    def __init__(self, dim=64, levels=7, center=(0, 0, 0), vmin=(-1, -1, -1), vmax=(1, 1, 1),
                 storage=abi.VCT_STORE_FP16, anisotropic=0, device=0, handle=None, owner=None):
        self.lib = load_library()
        self._owner = owner
        if handle is not None:
            self.h = C.c_void_p(handle)
            self._own = False
            return
        d = abi.GridDesc(dim, levels, (C.c_float * 3)(*center), (C.c_float * 3)(*vmin), (C.c_float * 3)(*vmax),
                         storage, anisotropic, device)
        h = C.c_void_p()
        rc = self.lib.vct_create(C.byref(h), C.byref(d))
        if rc:
            raise VCTError(rc, self.lib.vct_last_error(None).decode())
        self.h = h
        self._own = True

    def close(self):
        if getattr(self, "_own", False) and self.h:
            self.lib.vct_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int) -> None:
        if rc:
            raise VCTError(rc, (self.lib.vct_last_error(self.h) or b"").decode())

    # --- state ----------------------------------------------------------
    def grid(self) -> abi.GridDesc:
        d = abi.GridDesc()
        self.check(self.lib.vct_get_grid(self.h, C.byref(d)))
        return d

    def set_scene(self, pos, nrm, tan, albedo, obj, nobj, rough=None, vtan=None, uv=None, texid=None):
        """vct_set_scene; `vtan` [num_tris][3][3] per-corner tangents or None;
        `uv` [num_tris][3][2] texture coordinates and `texid` [num_tris]
        diffuse-map indices (-1: none) or None."""
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        tan = np.ascontiguousarray(tan, np.float32)
        albedo = np.ascontiguousarray(albedo, np.uint32)
        obj = np.ascontiguousarray(obj, np.uint32)
        rough_a = None if rough is None else np.ascontiguousarray(rough, np.float32)
        vtan_a = None if vtan is None else np.ascontiguousarray(vtan, np.float32)
        if vtan_a is not None and vtan_a.size != albedo.size * 9:
            raise ValueError("vtan must be [num_tris][3][3]")
        uv_a = None if uv is None else np.ascontiguousarray(uv, np.float32)
        tid_a = None if texid is None else np.ascontiguousarray(texid, np.int32)
        if tid_a is not None and (uv_a is None or uv_a.size != albedo.size * 6 or tid_a.size != albedo.size):
            raise ValueError("texid needs uv [num_tris][3][2] and one index per triangle")
        m = abi.Mesh(int(albedo.size), int(nobj), _ptr(pos), _ptr(nrm), _ptr(tan), _ptr(albedo), _ptr(obj),
                     None if rough_a is None else _ptr(rough_a), None if vtan_a is None else _ptr(vtan_a),
                     None if uv_a is None else _ptr(uv_a), None if tid_a is None else _ptr(tid_a))
        self.check(self.lib.vct_set_scene(self.h, C.byref(m)))

    def set_textures(self, textures):
        """vct_set_textures: a list of (width, height, levels, texels) with texels
        the uint32 RGBA8 words of every level back to back."""
        keep = [np.ascontiguousarray(t[3], np.uint32) for t in textures]
        arr = (abi.Texture * max(1, len(textures)))(*[abi.Texture(int(t[0]), int(t[1]), int(t[2]), _ptr(a))
                                                      for t, a in zip(textures, keep)])
        self.check(self.lib.vct_set_textures(self.h, C.cast(arr, C.c_void_p) if textures else None, len(textures)))

    def set_transforms(self, model: np.ndarray):
        model = np.ascontiguousarray(model, np.float32)
        self.check(self.lib.vct_set_transforms(self.h, _ptr(model), model.size // 16))

    def set_lights(self, lights):
        arr = (abi.Light * len(lights))(*lights)
        self.check(self.lib.vct_set_lights(self.h, arr, len(lights)))

    # --- passes (Application::render order) ---------------------------------
    def shadowmap(self, i):
        self.check(self.lib.vct_shadowmap(self.h, i))

    def shadowmaps(self):
        self.check(self.lib.vct_shadowmaps(self.h))

    def voxelize(self, p: abi.VoxelizeParams):
        self.check(self.lib.vct_voxelize(self.h, C.byref(p)))

    def voxelize_tessellated(self, p: abi.VoxelizeParams):
        self.check(self.lib.vct_voxelize_tessellated(self.h, C.byref(p)))

    def normalize_voxels(self, opacity: float = 0.5):
        """normalizeVoxels.comp after a VCT_ACCUM_SUM voxelization."""
        self.check(self.lib.vct_normalize_voxels(self.h, opacity))

    def normalize_inject(self, p: abi.InjectParams):
        self.check(self.lib.vct_normalize_inject(self.h, C.byref(p)))

    def normalize_inject_noalloc(self, p: abi.InjectParams):
        """The same pass with its scratch in the library's static device array
        (no allocation; grids up to 256^3)."""
        self.check(self.lib.vct_normalize_inject_noalloc(self.h, C.byref(p)))

    def fill_holes(self):
        self.check(self.lib.vct_fill_holes(self.h))

    def mip(self, p: abi.MipParams | None = None):
        p = p or abi.MipParams(0, 0)
        self.check(self.lib.vct_mip(self.h, C.byref(p)))

    def gbuffer(self, cam: abi.Camera):
        self.check(self.lib.vct_gbuffer(self.h, C.byref(cam)))

    def trace(self, p: abi.TraceParams):
        self.check(self.lib.vct_trace(self.h, C.byref(p)))

    def warpmap(self, vox: abi.VoxelizeParams, p: abi.WarpmapParams | None = None):
        self.check(self.lib.vct_warpmap(self.h, C.byref(vox), C.byref(p or abi.default_warpmap_params())))

    def warp_texture(self) -> np.ndarray:
        """The 32^3 warp texture (z, y, x, 4): warped coordinate xyz, packCellInfo w."""
        return self.download(abi.VCT_BUF_WARPMAP, 0, np.float32).reshape(32, 32, 32, 4)

    def warp_occupancy(self) -> np.ndarray:
        return self.download(abi.VCT_BUF_WARP_OCCUPANCY, 0, np.uint32).reshape(32, 32, 32)

    def compose(self, p: abi.ComposeParams):
        self.check(self.lib.vct_compose(self.h, C.byref(p)))

    def color(self) -> np.ndarray:
        """Final image of the last vct_compose, float4 per pixel."""
        return self.download(abi.VCT_BUF_COLOR, 0, np.float32).reshape(-1, 4)

    def sync(self):
        self.check(self.lib.vct_sync(self.h))

    def set_timing(self, on: bool):
        self.check(self.lib.vct_set_timing(self.h, 1 if on else 0))

    def stats(self) -> abi.Stats:
        s = abi.Stats()
        self.check(self.lib.vct_get_stats(self.h, C.byref(s)))
        return s

    def pass_times(self) -> abi.PassTimes:
        t = abi.PassTimes()
        self.check(self.lib.vct_get_pass_times(self.h, C.byref(t)))
        return t

    # --- readback ----------------------------------------------------------
    def nbytes(self, what: int, level: int = 0) -> int:
        return int(self.lib.vct_buffer_bytes(self.h, what, level))

    def download(self, what: int, level: int = 0, dtype=np.uint8) -> np.ndarray:
        n = self.nbytes(what, level)
        if n == 0:
            raise VCTError(abi.VCT_ERR_INVALID, f"buffer {what} level {level} does not exist")
        out = np.empty(n, np.uint8)
        self.check(self.lib.vct_download(self.h, what, level, _ptr(out), n))
        return out.view(dtype)

    def upload(self, what: int, level: int, data: np.ndarray):
        data = np.ascontiguousarray(data)
        self.check(self.lib.vct_upload(self.h, what, level, _ptr(data), data.nbytes))

    def albedo(self) -> np.ndarray:
        d = self.grid().dim
        return self.download(abi.VCT_BUF_ALBEDO, 0, np.uint32).reshape(d, d, d)

    def normal(self) -> np.ndarray:
        d = self.grid().dim
        return self.download(abi.VCT_BUF_NORMAL, 0, np.uint32).reshape(d, d, d)

    def radiance(self, level: int = 0, face: int = 0) -> np.ndarray:
        d = self.grid().dim
        s = max(1, d >> level)
        return self.download(abi.VCT_BUF_RADIANCE, level | (face << 4), np.float32).reshape(s, s, s, 4)

    def shadow(self, light: int) -> np.ndarray:
        a = self.download(abi.VCT_BUF_SHADOWMAP, light, np.float32)
        n = int(round(a.size ** 0.5))
        return a.reshape(n, n)

    def litmask(self) -> np.ndarray:
        return self.download(abi.VCT_BUF_LITMASK, 0, np.uint8)

    # --- learned diffuse gather ------------------------------------------------
    def learned_set_points(self, points: np.ndarray):
        points = np.ascontiguousarray(points, np.float32).reshape(-1, 16)
        self.check(self.lib.vct_learned_set_points(self.h, _ptr(points), points.shape[0]))

    def learned_prepare(self):
        self.check(self.lib.vct_learned_prepare(self.h))

    def learned_train(self, steps=20000, batch=4096, lr=1e-3, seed=1, reset=False, lr_final=0.0) -> np.ndarray:
        p = abi.LearnedParams(batch, steps, lr, 0.9, 0.999, 1e-8, seed, 1 if reset else 0, lr_final)
        loss = np.zeros(max(steps, 1), np.float32)
        self.check(self.lib.vct_learned_train(self.h, C.byref(p), _ptr(loss)))
        return loss[:steps]

    def trace_learned(self, p: abi.TraceParams, threshold: float, count_fallback=True) -> int:
        n = C.c_int(0)
        self.check(self.lib.vct_trace_learned(self.h, C.byref(p), threshold, C.byref(n) if count_fallback else None))
        return n.value

    def reference_gather(self):
        self.check(self.lib.vct_reference_gather(self.h))

    def weights(self) -> np.ndarray:
        return self.download(abi.VCT_BUF_WEIGHTS, 0, np.float32)

    def set_weights(self, w: np.ndarray):
        self.upload(abi.VCT_BUF_WEIGHTS, 0, np.ascontiguousarray(w, np.float32))


def load_obj(path) -> dict:
    """An OBJ/MTL file as the triangle-soup arrays of vct_mesh (one object),
    plus its bounds and material count (vct_obj_load)."""
    lib = load_app_library()
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = lib.vct_obj_load(str(path).encode(), C.byref(h), err, 256)
    if rc:
        raise VCTError(rc, err.value.decode())
    try:
        m = abi.Mesh()
        lib.vct_obj_mesh(h, C.byref(m))
        n = m.num_tris

        def arr(ptr, count, dt):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                         shape=(count,)).copy()

        bmin, bmax = np.zeros(3, np.float32), np.zeros(3, np.float32)
        nm = C.c_int()
        lib.vct_obj_bounds(h, _ptr(bmin), _ptr(bmax), C.byref(nm))
        tp, nt = C.c_void_p(), C.c_int()
        lib.vct_obj_textures(h, C.byref(tp), C.byref(nt))
        textures = []
        for i in range(nt.value):
            t = abi.Texture.from_address(tp.value + i * C.sizeof(abi.Texture))
            textures.append(_texture_tuple(t))
        return dict(pos=arr(m.positions, n * 9, np.float32).reshape(n, 3, 3),
                    nrm=arr(m.normals, n * 9, np.float32).reshape(n, 3, 3),
                    tan=arr(m.tangents, n * 3, np.float32).reshape(n, 3),
                    vtan=None if not m.vertex_tangents else arr(m.vertex_tangents, n * 9, np.float32).reshape(n, 3, 3),
                    albedo=arr(m.albedo, n, np.uint32), obj=arr(m.object, n, np.uint32), nobj=1,
                    rough=None if not m.roughness else arr(m.roughness, 1, np.float32),
                    bmin=bmin, bmax=bmax, materials=nm.value,
                    uv=arr(m.texcoords, n * 6, np.float32).reshape(n, 3, 2),
                    texid=arr(m.texture, n, np.int32), textures=textures)
    finally:
        lib.vct_obj_free(h)


def _texture_tuple(t) -> tuple:
    """(width, height, levels, texels) of a vct_texture view, texels copied."""
    n = sum(max(1, t.width >> l) * max(1, t.height >> l) for l in range(t.levels))
    if n == 0:
        return (t.width, t.height, t.levels, np.zeros(0, np.uint32))
    a = np.ctypeslib.as_array(C.cast(t.texels, C.POINTER(C.c_uint32)), shape=(n,)).copy()
    return (t.width, t.height, t.levels, a)


def load_dds(path) -> tuple:
    """A DDS file (DXT1/3/5) decoded to RGBA8 levels (vct_dds_load):
    (width, height, levels, texels of every level back to back)."""
    return _load_image(path, "vct_dds_load")


def load_png(path) -> tuple:
    """A PNG file as the reference uploads it (vct_png_load): (width, height,
    levels, texels of every level back to back); 0 levels for two channels."""
    return _load_image(path, "vct_png_load")


def _load_image(path, fn) -> tuple:
    lib = load_app_library()
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = getattr(lib, fn)(str(path).encode(), C.byref(h), err, 256)
    if rc:
        raise VCTError(rc, err.value.decode())
    try:
        t = abi.Texture()
        lib.vct_dds_texture(h, C.byref(t))
        return _texture_tuple(t)
    finally:
        lib.vct_dds_free(h)


class App:
    """Owns one vct_app (the headless Application of include/vct_app.h)."""

    def __init__(self, scene=1, obj_path=None, **overrides):
        self.lib = load_app_library()
        cfg = abi.AppConfig()
        rc = self.lib.vct_app_default_config(scene if obj_path is None else 1, C.byref(cfg))
        if rc:
            raise VCTError(rc, f"unknown scene {scene}")
        for k, v in overrides.items():
            setattr(cfg, k, v)
        h = C.c_void_p()
        if obj_path is None:
            rc = self.lib.vct_app_create(C.byref(h), C.byref(cfg))
        else:
            rc = self.lib.vct_app_create_from_obj(C.byref(h), C.byref(cfg), str(obj_path).encode())
        if rc:
            raise VCTError(rc, (self.lib.vct_app_last_error(None) or b"").decode())
        self.h = h
        ctx_handle = self.lib.vct_app_ctx(h)
        self.ctx = Context(handle=ctx_handle, owner=self) if ctx_handle else None
        self.cfg = abi.AppConfig()
        self.lib.vct_app_get_config(h, C.byref(self.cfg))

    def close(self):
        if getattr(self, "h", None):
            if self.ctx is not None:
                self.ctx.h = None
            self.lib.vct_app_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int):
        if rc:
            raise VCTError(rc, (self.lib.vct_app_last_error(self.h) or b"").decode())

    def sizes(self):
        nt, no, nl = C.c_int(), C.c_int(), C.c_int()
        self.check(self.lib.vct_app_scene_sizes(self.h, C.byref(nt), C.byref(no), C.byref(nl)))
        return nt.value, no.value, nl.value

    def mesh(self) -> dict:
        nt, no, _ = self.sizes()
        pos = np.empty((nt, 3, 3), np.float32)
        nrm = np.empty((nt, 3, 3), np.float32)
        tan = np.empty((nt, 3), np.float32)
        alb = np.empty(nt, np.uint32)
        obj = np.empty(nt, np.uint32)
        rough = np.empty(no, np.float32)
        has = self.lib.vct_app_get_mesh(self.h, _ptr(pos), _ptr(nrm), _ptr(tan), _ptr(alb), _ptr(obj), _ptr(rough))
        vtan = np.empty((nt, 3, 3), np.float32)
        has_vtan = self.lib.vct_app_get_vertex_tangents(self.h, _ptr(vtan))
        return dict(pos=pos, nrm=nrm, tan=tan, albedo=alb, obj=obj, nobj=no, rough=rough if has == 1 else None,
                    vtan=vtan if has_vtan == 1 else None)

    def frame_state(self, frame: int = 0) -> dict:
        _, no, nl = self.sizes()
        model = np.empty((no, 16), np.float32)
        lights = (abi.Light * nl)()
        vox, inj, mip = abi.VoxelizeParams(), abi.InjectParams(), abi.MipParams()
        cam, tr = abi.Camera(), abi.TraceParams()
        self.check(self.lib.vct_app_frame_state(self.h, frame, _ptr(model), lights, C.byref(vox), C.byref(inj),
                                                C.byref(mip), C.byref(cam), C.byref(tr)))
        return dict(model=model, lights=list(lights), vox=vox, inj=inj, mip=mip, cam=cam, trace=tr)

    def compose_params(self, frame: int = 0) -> abi.ComposeParams:
        p = abi.ComposeParams()
        self.check(self.lib.vct_app_compose_params(self.h, frame, C.byref(p)))
        return p

    def surface_points(self, n: int, frame: int = 0, seed: int = 1) -> np.ndarray:
        out = np.zeros((n, 16), np.float32)
        self.check(self.lib.vct_app_surface_points(self.h, frame, n, seed, _ptr(out)))
        return out

    def camera(self, seed: int, width: int | None = None, height: int | None = None) -> abi.Camera:
        """A seeded camera of the scene's camera distribution."""
        cam = abi.Camera()
        w = width or self.cfg.width
        h = height or self.cfg.height
        self.check(self.lib.vct_app_camera(self.h, seed, w, h, C.byref(cam)))
        return cam

    def render(self, frame: int = 0):
        self.check(self.lib.vct_app_render(self.h, frame))

    def build_radiance(self, frame: int = 0):
        self.check(self.lib.vct_app_build_radiance(self.h, frame))
