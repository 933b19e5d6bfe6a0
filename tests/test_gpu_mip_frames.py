"""GPU parity of the radiance mip chain over consecutive frames
(filterRadiance.comp KERNEL_BOX2, Application.cpp:877-921, once per frame).

The chain does not read level-0 bricks that the injection recorded as +0, a
record carried from frame to frame.  The scene alternates between its own
placement and one shifted by a third of the grid, so 8^3 regions go from
empty to filled and back while others stay empty; every frame's levels and
faces are bit-exact against the oracle's chain of that frame's level 0, and
the albedo chain against the oracle's chain of that frame's albedo grid (the
injection records the grid's all-zero bricks, which that chain does not read;
a voxelization without an injection after it voids the record).
"""
import numpy as np
import pytest

from gpu_util import make_app
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


def check_albedo_chain(ctx, g, alb, label):
    """The voxelColor chain of the same mipmap pass (Application.cpp:904-917)
    against the oracle's, from the albedo grid as the chain read it."""
    ref = ob.mip_albedo(alb, g.levels)
    for l in range(1, g.levels):
        got = ctx.download(abi.VCT_BUF_ALBEDO, l, np.uint32).reshape((max(1, g.dim >> l),) * 3)
        assert np.array_equal(got, ref[l]), (label, l, int((got != ref[l]).sum()))


@pytest.mark.parametrize("aniso", [1, 0], ids=["aniso", "iso"])
def test_mip_chain_bit_exact_over_frames(aniso):
    app = make_app(2, dim=128, levels=7, width=256, height=144, shadow_size=256, anisotropic=aniso)
    try:
        ctx = app.ctx
        g = ctx.grid()
        shift = (g.vmax[0] - g.vmin[0]) / 3.0
        empty = []
        for frame in range(5):
            fs = app.frame_state(frame)
            model = fs["model"].copy()
            if frame % 2 == 1:
                model[:, 12] += shift  # column-major: the translation's x
            print(f"mip frames (aniso {aniso}): frame {frame}", flush=True)
            ctx.set_transforms(model)
            for i, L in enumerate(fs["lights"]):
                if L.shadow_caster:
                    ctx.shadowmap(i)
            ctx.voxelize(fs["vox"])
            ctx.normalize_inject(fs["inj"])
            rad0 = ctx.radiance(0)
            alb = ctx.albedo()
            ctx.mip(fs["mip"])
            check_albedo_chain(ctx, g, alb, frame)
            occupied = np.abs(rad0.reshape(16, 8, 16, 8, 16, 8, 4)).sum(axis=(1, 3, 5, 6)) > 0
            empty.append(~occupied)
            assert occupied.any()
            if aniso:
                ref = ob.mip_aniso(rad0, g.levels)
                for l in range(1, g.levels):
                    for f in range(6):
                        got = ctx.radiance(l, f)
                        assert np.array_equal(got.view(np.uint32), ref[l][f].view(np.uint32)), (frame, l, f)
            else:
                ref = ob.mip(rad0, g.levels)
                for l in range(1, g.levels):
                    got = ctx.radiance(l)
                    assert np.array_equal(got.view(np.uint32), ref[l].view(np.uint32)), (frame, l)
        # a voxelization followed by the chain without an injection: the
        # chain reads the new grid whole (the injection's record of empty
        # bricks describes the grid it saw, not this one)
        fs = app.frame_state(0)
        model = fs["model"].copy()
        model[:, 12] -= shift
        ctx.set_transforms(model)
        ctx.voxelize(fs["vox"])
        ctx.mip(fs["mip"])
        check_albedo_chain(ctx, g, ctx.albedo(), "voxelize + mip")
        # the frames exercised every transition of a region
        e = np.stack(empty)
        assert (e[:-1] & e[1:]).any()      # empty in two frames in a row
        assert (e[:-1] & ~e[1:]).any()     # filled after being empty
        assert (~e[:-1] & e[1:]).any()     # emptied after being filled
    finally:
        app.close()


def test_voxel_grids_bit_exact_over_frames():
    """The voxelize pass clears both voxel grids (Application.cpp:686-687)
    except the bricks the last injection found empty behind the standard
    voxelizer; the grids of every frame are bit-exact against the oracle's
    voxelization, with bricks emptied and refilled between frames, after a
    tessellated voxelization (whose albedo words can be 0 where it wrote a
    normal, so the next clear is whole) and after a second injection of the
    same grids."""
    app = make_app(2, dim=128, levels=7, width=256, height=144, shadow_size=256)
    try:
        ctx = app.ctx
        g = ctx.grid()
        shift = (g.vmax[0] - g.vmin[0]) / 3.0
        mesh = app.mesh()

        def frame(k, dx, tess=False, injections=1):
            fs = app.frame_state(k)
            model = fs["model"].copy()
            model[:, 12] += dx
            ctx.set_transforms(model)
            sms = []
            for i, L in enumerate(fs["lights"]):
                if L.shadow_caster:
                    ctx.shadowmap(i)
                    sms.append(ctx.shadow(i))
                else:
                    sms.append(None)
            sc = ob.SceneArrays(mesh, model)
            vox = fs["vox"]
            vox.atomic_max = abi.VCT_ACCUM_MAX
            print(f"voxel grids: frame {k}, shift {dx:.3f}, tessellated {tess}", flush=True)
            if tess:
                ctx.voxelize_tessellated(vox)
                ralb, rnrm = ob.voxelize_tess(sc, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax), accum=1)
            else:
                ctx.voxelize(vox)
                ralb, rnrm, _ = ob.voxelize(sc, fs["lights"], sms, vox, g.dim)
            alb, nrm = ctx.albedo(), ctx.normal()
            assert np.array_equal(alb, ralb), (k, int((alb != ralb).sum()))
            assert np.array_equal(nrm, rnrm), (k, int((nrm != rnrm).sum()))
            for _ in range(injections):
                ctx.normalize_inject(fs["inj"])
            return (alb != 0).reshape(16, 8, 16, 8, 16, 8).any(axis=(1, 3, 5))

        occ = [frame(k, shift * (k % 2)) for k in range(4)]
        occ.append(frame(4, 0.0, tess=True))
        occ.append(frame(5, shift))
        occ.append(frame(6, 0.0, injections=2))
        occ.append(frame(7, shift))
        e = ~np.stack(occ)
        assert (e[:-1] & ~e[1:]).any() and (~e[:-1] & e[1:]).any()  # regions filled and emptied
    finally:
        app.close()
