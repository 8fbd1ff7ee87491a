"""CPU tests of the native library: ABI surface and host-side scene logic
(flattening, edge lists, BVH invariants).  No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from tensorray_amd import _capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            src = open(os.path.join(inc, fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(trc_[a-z_0-9]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = _capi.load_library()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    bound = {s[0] for s in _capi.SIGNATURES}
    assert declared <= bound, declared - bound


@pytest.fixture(scope="module")
def host_ctx():
    ctx = _capi.Context(host_only=True)
    yield ctx
    ctx.close()


def _edges_reference(idx):
    """ObjMesh::processEdges (EDXUtil/Graphics/ObjMesh.cpp:584-632)."""
    m = {}
    for i, f in enumerate(idx.tolist()):
        for j in range(3):
            a, b, c = f[j], f[(j + 1) % 3], f[(j + 2) % 3]
            key = (min(a, b), max(a, b))
            if key not in m:
                m[key] = [c]
            m[key].append(i)
    out = []
    for key in sorted(m):
        v = m[key]
        out.append([key[0], key[1], v[0], v[1], v[2] if len(v) == 3 else -1])
    return np.array(out, np.int32)


def test_edge_list_semantics(host_ctx):
    spec = scenes.sphere_scene(subdiv=2)
    ds = _capi.DeviceScene(host_ctx, spec)
    got = ds.edges()
    voff = toff = 0
    ref = []
    for m in spec.meshes:
        e = _edges_reference(np.asarray(m.idx))
        e[:, :3] += voff
        e[:, 3] += toff
        e[:, 4] = np.where(e[:, 4] >= 0, e[:, 4] + toff, -1)
        ref.append(e)
        voff += len(m.pos)
        toff += len(m.idx)
    ref = np.concatenate(ref)
    assert np.array_equal(got, ref)
    # closed icosphere: every edge has two faces; the quads have boundary edges
    n_sphere = len(_edges_reference(np.asarray(spec.meshes[0].idx)))
    assert (got[:n_sphere, 4] >= 0).all()
    assert (got[n_sphere:, 4] == -1).any()
    ds.close()


@pytest.mark.parametrize("subdiv,amp", [(2, 0.0), (4, 0.05)])
def test_bvh_invariants(host_ctx, subdiv, amp):
    spec = scenes.sphere_scene(subdiv=subdiv, amplitude=amp)
    ds = _capi.DeviceScene(host_ctx, spec)
    info = ds.info()
    nodes, order = ds.bvh()
    T = info.num_triangles
    assert sorted(order.tolist()) == list(range(T))  # each triangle exactly once
    f = spec.flat()
    pos, idx = f["pos"], f["idx"]
    fl = nodes.view(np.float32)
    child = nodes[:, 24:28]
    seen = np.zeros(T, np.int64)
    depth_max = 0
    stack = [(0, 1, None)]
    while stack:
        n, depth, parent_box = stack.pop()
        depth_max = max(depth_max, depth)
        for s in range(4):
            c = int(child[n, s])
            if c == 0xFFFFFFFF:
                continue
            lo = fl[n, [s, 4 + s, 8 + s]]
            hi = fl[n, [12 + s, 16 + s, 20 + s]]
            if parent_box is not None:
                assert (lo >= parent_box[0] - 1e-5).all() and (hi <= parent_box[1] + 1e-5).all()
            if c & 0x80000000:
                first, cnt = c & ((1 << 27) - 1), ((c >> 27) & 0xF) + 1
                for t in order[first:first + cnt]:
                    seen[t] += 1
                    v = pos[idx[t]]
                    assert (v >= lo).all() and (v <= hi).all()  # leaf box contains its triangles
            else:
                assert c > n  # children after parents (refit order)
                stack.append((c, depth + 1, (lo, hi)))
    assert (seen == 1).all()
    assert depth_max <= info.bvh_max_depth + 1
    assert info.bvh_stack_need <= 3 * info.bvh_max_depth + 1
    ds.close()


def test_invalid_scene_is_rejected(host_ctx):
    spec = scenes.config1_scene()
    spec.meshes[0].idx = spec.meshes[0].idx.copy()
    spec.meshes[0].idx[0, 0] = 10 ** 6
    with pytest.raises(_capi.TRCError, match="index out of range"):
        _capi.DeviceScene(host_ctx, spec)


def test_host_context_cannot_render(host_ctx):
    ds = _capi.DeviceScene(host_ctx, scenes.config1_scene())
    cam = _capi.make_camera(8, 8)
    opt = _capi.make_options(spp=1)
    with pytest.raises(_capi.TRCError, match="TRC_ERR_STATE"):
        ds.render_fwd(cam, opt, 0)
    ds.close()


def test_host_context_rejects_device_work(host_ctx):
    """Batches, plans, optimizer steps and solvers need a device; on a
    host-only context they fail with TRC_ERR_STATE before touching memory."""
    ds = _capi.DeviceScene(host_ctx, scenes.config1_scene())
    cam = _capi.make_camera(8, 8)
    opt = _capi.make_options(spp=1)
    job = ds.job(cam, opt, 0, 64, paths=1, ckpt=1)
    with pytest.raises(_capi.TRCError, match="TRC_ERR_STATE"):
        host_ctx.render_batch([job], False)
    with pytest.raises(_capi.TRCError, match="TRC_ERR_STATE"):
        host_ctx.plan([job], adjoint=False)
    with pytest.raises(_capi.TRCError, match="TRC_ERR_STATE"):
        host_ctx.adam_step(1, 1, 1, 1, 4, 1)
    with pytest.raises(_capi.TRCError, match="TRC_ERR_STATE"):
        host_ctx.adam_uniform_step(1, 1, 1, 1, 4, 1)
    with pytest.raises(_capi.TRCError, match="TRC_ERR_STATE"):
        host_ctx.lsolver(np.array([[0, 1, 2]]), 3, 1.0)
    for kind in ("interior", "primary"):
        dopt = _capi.make_options(spp=1, sppe=1)
        with pytest.raises(_capi.TRCError, match="TRC_ERR_STATE"):
            ds.render_deriv(cam, dopt, kind, 1, None, 1)
        with pytest.raises(_capi.TRCError, match="TRC_ERR_INVALID"):
            ds.render_deriv(cam, dopt, kind, None, None, 1)   # no tangent buffer
    with pytest.raises(_capi.TRCError, match="TRC_ERR_INVALID"):
        ds.render_deriv(cam, opt, 7, 1, None, 1)              # no such integrator
    ds.close()
