"""GPU forward/adjoint pair (differentiable depth/shading) on the 81,920-face
icosphere, 8 cameras at 512x512: the forward is bitwise the oracle's binary32
forward; the adjoint matches the oracle's binary64 gradient to 1e-4 relative
and central finite differences of the GPU's own forward on pixels whose hit
triangle does not change."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import scenes
from helpers import assert_hits_equal

pytestmark = pytest.mark.gpu

W = 512
NCAM = 8
NRAY = NCAM * W * W


def gpu_rays(ctx, cams, width):
    n1 = width * width
    d = ctx.alloc(len(cams) * n1 * B.RAY_DTYPE.itemsize)
    for k in range(len(cams)):
        view = B.DeviceArray.view(ctx, d.value + k * n1 * B.RAY_DTYPE.itemsize, n1 * B.RAY_DTYPE.itemsize)
        B.raygen(ctx, cams[k:k + 1], width, width, B.ORDER_TILES, d_rays=view)
    return d


def make_mesh(ctx, V, F):
    soup = np.ascontiguousarray(V[F].reshape(-1, 9))
    orig = soup.copy()
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(soup)
    return B.Mesh(bvh, V, F), orig


@pytest.fixture(scope="module")
def scene(ctx):
    w = scenes.diff_workload(64, width=W, n_cameras=NCAM)
    d_rays = gpu_rays(ctx, w["cameras"], W)
    target_mesh, _ = make_mesh(ctx, w["V_target"], w["F"])
    d_hits = ctx.alloc(NRAY * B.HIT_DTYPE.itemsize)
    d_target = ctx.alloc(NRAY * 8)
    B.trace_forward(target_mesh, w["Minv"], w["light"], d_rays, NRAY, d_hits, d_target)
    mesh, orig = make_mesh(ctx, w["V"], w["F"])
    ctx.sync()
    return dict(w=w, d_rays=d_rays, d_target=d_target, target_mesh=target_mesh, mesh=mesh, orig=orig,
                rays=d_rays.download(B.RAY_DTYPE, NRAY), target=d_target.download(np.float32, 2 * NRAY).reshape(-1, 2))


def forward(ctx, scene, mesh=None, minv=None):
    mesh = mesh or scene["mesh"]
    w = scene["w"]
    d_hits = ctx.alloc(NRAY * B.HIT_DTYPE.itemsize)
    d_img = ctx.alloc(NRAY * 8)
    B.trace_forward(mesh, w["Minv"] if minv is None else minv, w["light"], scene["d_rays"], NRAY, d_hits, d_img)
    return d_hits, d_img


def adjoint(ctx, scene, d_rays, n, d_hits, d_target, w_depth=1.0, w_shade=1.0, mesh=None):
    mesh = mesh or scene["mesh"]
    w = scene["w"]
    d_gv = ctx.alloc(mesh.n_vertices * 12)
    d_gm = ctx.alloc(64)
    d_loss = ctx.alloc(8)
    B.trace_adjoint(mesh, w["Minv"], w["light"], d_rays, n, d_hits, d_target, d_gv, d_gm, d_loss, w_depth, w_shade)
    return d_gv, d_gm, d_loss


def vid_of(mesh):
    nodes, used, tris = mesh.bvh.download()
    return np.ascontiguousarray(mesh.prim_vertices().reshape(-1)), nodes, used, tris


def test_permutation_and_binding(ctx, orc, scene):
    mesh, orig = scene["mesh"], scene["orig"]
    perm = mesh.bvh.permutation()
    assert np.array_equal(np.sort(perm), np.arange(orig.shape[0]))
    _, _, tris = mesh.bvh.download()
    assert tris.tobytes() == orig[perm].tobytes()
    V, N = mesh.download()
    assert V.tobytes() == scene["w"]["V"].tobytes()
    assert N.tobytes() == orc.mesh_normals(scene["w"]["V"], scene["w"]["F"]).tobytes()
    with pytest.raises(B.BvhtError, match="not the mesh faces"):
        B.Mesh(mesh.bvh, scene["w"]["V"], scene["w"]["F"], face_of=np.roll(perm, 1))
    with pytest.raises(B.BvhtError, match="out of range"):
        B.Mesh(mesh.bvh, scene["w"]["V"], scene["w"]["F"], face_of=np.full(perm.shape, 1 << 30, np.uint32))


@pytest.mark.parametrize("which", ["target", "start"])
def test_forward_bitwise_equals_oracle(ctx, orc, scene, which):
    mesh = scene["target_mesh"] if which == "target" else scene["mesh"]
    d_hits, d_img = forward(ctx, scene, mesh)
    hits = d_hits.download(B.HIT_DTYPE, NRAY)
    img = d_img.download(np.float32, 2 * NRAY).reshape(-1, 2)
    vid, nodes, _, tris = vid_of(mesh)
    _, N = mesh.download()
    w = scene["w"]
    ref_h, ref_img = orc.diff_forward(nodes, tris, vid, N, w["Minv"], w["light"], scene["rays"].view(orc.RAY_DTYPE),
                                      nthreads=16)
    assert_hits_equal(hits, ref_h, f"forward hits ({which})")
    assert img.tobytes() == ref_img.tobytes()
    assert (hits["prim"] != B.MISS).sum() > NRAY // 3


def test_forward_wave_schedule_keeps_results(ctx, scene):
    """Each forward orders its 64-ray groups longest first from the durations
    the previous forward over the same ray count recorded.  The results are
    the unscheduled forward's bit for bit, on repeated calls, after a call
    with another ray count, and after the mesh moved (SGD step + refit)."""
    w = scene["w"]
    mesh, _ = make_mesh(ctx, w["V"], w["F"])
    first = [a.download(np.uint8).tobytes() for a in forward(ctx, scene, mesh)]
    for _ in range(2):
        again = [a.download(np.uint8).tobytes() for a in forward(ctx, scene, mesh)]
        assert again == first
    half = NRAY // 2 + 37
    d_h, d_i = ctx.alloc(half * B.HIT_DTYPE.itemsize), ctx.alloc(half * 8)
    B.trace_forward(mesh, w["Minv"], w["light"], scene["d_rays"], half, d_h, d_i)
    B.trace_forward(mesh, w["Minv"], w["light"], scene["d_rays"], half, d_h, d_i)  # scheduled, odd count
    assert d_h.download(np.uint8).tobytes() == first[0][:half * B.HIT_DTYPE.itemsize]
    assert d_i.download(np.uint8).tobytes() == first[1][:half * 8]
    assert [a.download(np.uint8).tobytes() for a in forward(ctx, scene, mesh)] == first
    # one gradient step applied to this mesh and to a fresh twin without a
    # schedule: the scheduled forward of the moved mesh is the twin's
    d_hits, _ = forward(ctx, scene, mesh)
    d_gv, _, _ = adjoint(ctx, scene, scene["d_rays"], NRAY, d_hits, scene["d_target"], mesh=mesh)
    twin, _ = make_mesh(ctx, w["V"], w["F"])
    for m in (mesh, twin):
        m.sgd(d_gv, scenes.DIFF_LR, scenes.DIFF_MAX_STEP)
        m.update(normals=True)
    assert [a.download(np.uint8).tobytes() for a in forward(ctx, scene, mesh)] == \
        [a.download(np.uint8).tobytes() for a in forward(ctx, scene, twin)]


def test_forward_schedule_of_another_count_between_graph_replays(ctx, scene):
    """A graph captures a scheduled forward of NRAY rays; between its
    replays an eager forward of another ray count rebuilds the order for its
    own count.  A replay then runs its groups in slot order (the order's
    recorded unit count differs from its own), and every replay equals the
    forward made eagerly, bit for bit."""
    w = scene["w"]
    mesh, _ = make_mesh(ctx, w["V"], w["F"])
    want = [a.download(np.uint8).tobytes() for a in forward(ctx, scene, mesh)]
    forward(ctx, scene, mesh)  # the order now exists for NRAY rays
    d_hits, d_img = ctx.alloc(NRAY * B.HIT_DTYPE.itemsize), ctx.alloc(NRAY * 8)

    def enqueue():
        B.trace_forward(mesh, w["Minv"], w["light"], scene["d_rays"], NRAY, d_hits, d_img)

    g = B.Graph.capture(ctx, enqueue, keep=(d_hits, d_img))
    small = NRAY // 4 + 5
    d_h, d_i = ctx.alloc(small * B.HIT_DTYPE.itemsize), ctx.alloc(small * 8)
    try:
        for _ in range(3):
            B.trace_forward(mesh, w["Minv"], w["light"], scene["d_rays"], small, d_h, d_i)
            B.trace_forward(mesh, w["Minv"], w["light"], scene["d_rays"], small, d_h, d_i)
            d_hits.upload(np.zeros(NRAY, B.HIT_DTYPE))
            g.launch()
            ctx.sync()
            assert [d_hits.download(np.uint8).tobytes(), d_img.download(np.uint8).tobytes()] == want
            assert d_h.download(np.uint8).tobytes() == want[0][:small * B.HIT_DTYPE.itemsize]
    finally:
        g.close()


def gradient_errors(gv, gm, ref_gv, ref_gm):
    """Per-vertex error on the vertex's own scale, |g - g64| / max(|g64|,
    1e-6 max |g64|) (vector norms per vertex; the floor only matters for
    vertices whose gradient is a millionth of the largest), the normwise
    error ||g - g64|| / ||g64||, and the per-entry error of dLoss/dM^-1 on
    each entry's own scale."""
    mag = np.linalg.norm(ref_gv, axis=1)
    per_vertex = np.linalg.norm(gv - ref_gv, axis=1) / np.maximum(mag, 1e-6 * mag.max())
    normwise = np.linalg.norm(gv - ref_gv) / np.linalg.norm(ref_gv)
    nz = np.abs(ref_gm) > 0
    per_entry = np.abs(gm - ref_gm)[nz] / np.abs(ref_gm)[nz]
    return per_vertex, normwise, per_entry


@pytest.mark.parametrize("w_depth,w_shade", [(1.0, 1.0), (0.0, 1.0)])
def test_adjoint_matches_fp64_gradient(ctx, orc, scene, w_depth, w_shade):
    """Every vertex within 1e-4 of the binary64 gradient on its own scale,
    not only on the scale of the largest (silhouette) gradients; every
    dLoss/dM^-1 entry within 1e-4 of its own value."""
    mesh, w = scene["mesh"], scene["w"]
    d_hits, _ = forward(ctx, scene)
    d_gv, d_gm, d_loss = adjoint(ctx, scene, scene["d_rays"], NRAY, d_hits, scene["d_target"], w_depth, w_shade)
    gv = d_gv.download(np.float32).reshape(-1, 3)
    gm = d_gm.download(np.float32, 16)
    loss = float(d_loss.download(np.float64, 1)[0])
    hits = d_hits.download(B.HIT_DTYPE, NRAY)
    vid, _, _, _ = vid_of(mesh)
    _, N = mesh.download()
    ref_loss, ref_gv, ref_gm, _ = orc.diff_grad64(w["V"], vid, N, w["Minv"], w["light"],
                                                  scene["rays"].view(orc.RAY_DTYPE), hits["prim"], scene["target"],
                                                  w_depth, w_shade)
    assert ref_loss > 0
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    per_vertex, normwise, per_entry = gradient_errors(gv, gm, ref_gv, ref_gm)
    assert (np.linalg.norm(ref_gv, axis=1) > 0).sum() > 20000
    assert per_vertex.max() <= 1e-4, (per_vertex.max(), np.argmax(per_vertex))
    assert normwise <= 1e-6, normwise
    assert per_entry.max() <= 1e-4, per_entry
    assert np.all(gm[3::4] == 0)


def test_adjoint_is_bitwise_repeatable(ctx, scene):
    """The vertex scatter has no floating-point atomics: three adjoints of
    the same hits give the same bits, also with a forward (which reorders
    its waves) between them."""
    d_hits, _ = forward(ctx, scene)
    runs = []
    for k in range(3):
        d_gv, d_gm, d_loss = adjoint(ctx, scene, scene["d_rays"], NRAY, d_hits, scene["d_target"])
        runs.append((d_gv.download(np.uint8).tobytes(), d_gm.download(np.uint8).tobytes(),
                     d_loss.download(np.uint8).tobytes()))
        if k == 1:
            forward(ctx, scene)
    assert runs[1] == runs[0] and runs[2] == runs[0]


def test_adjoint_accumulate_and_zero_residual(ctx, scene):
    w = scene["w"]
    mesh = scene["mesh"]
    d_hits, _ = forward(ctx, scene)
    d_gv, d_gm, d_loss = adjoint(ctx, scene, scene["d_rays"], NRAY, d_hits, scene["d_target"])
    g1, m1, l1 = d_gv.download(np.float32), d_gm.download(np.float32), d_loss.download(np.float64)
    B.trace_adjoint(mesh, w["Minv"], w["light"], scene["d_rays"], NRAY, d_hits, scene["d_target"], d_gv, d_gm,
                    d_loss, accumulate=True)
    g2, m2, l2 = d_gv.download(np.float32), d_gm.download(np.float32), d_loss.download(np.float64)
    # fixed-order sums: the second call adds exactly the first call's values
    assert np.array_equal(m2, 2 * m1) and np.array_equal(l2, 2 * l1) and np.array_equal(g2, 2 * g1)
    # the target mesh against its own image: only the binary32 rounding of
    # the image's depth (the adjoint re-solves in binary64) is left
    tm = scene["target_mesh"]
    d_th, d_timg = forward(ctx, scene, tm)
    d_gv, d_gm, d_loss = adjoint(ctx, scene, scene["d_rays"], NRAY, d_th, d_timg, mesh=tm)
    assert d_loss.download(np.float64, 1)[0] < 1e-6 * l1[0]
    assert np.linalg.norm(d_gv.download(np.float32)) < 1e-2 * np.linalg.norm(g1)
    assert np.linalg.norm(d_gm.download(np.float32)) < 1e-2 * np.linalg.norm(m1)


def test_adjoint_of_a_coarse_mesh_with_large_triangles(ctx, orc):
    """An icosphere of 180 faces seen at 512x512 from 2 cameras: each face
    covers thousands of pixels, so its wave records (one per wave that hits
    it) overflow the per-thread sort and take the workgroup sort.  The
    gradient equals the binary64 one on every vertex's own scale, bit for
    bit on a repeat."""
    w = scenes.diff_workload(3, width=W, n_cameras=2)
    rays = gpu_rays(ctx, w["cameras"], W)
    n = 2 * W * W
    tm, _ = make_mesh(ctx, w["V_target"], w["F"])
    d_hits, d_target = ctx.alloc(n * B.HIT_DTYPE.itemsize), ctx.alloc(n * 8)
    B.trace_forward(tm, w["Minv"], w["light"], rays, n, d_hits, d_target)
    mesh, _ = make_mesh(ctx, w["V"], w["F"])
    B.trace_forward(mesh, w["Minv"], w["light"], rays, n, d_hits, ctx.alloc(n * 8))
    outs = []
    for _ in range(2):
        d_gv, d_gm, d_loss = ctx.alloc(mesh.n_vertices * 12), ctx.alloc(64), ctx.alloc(8)
        B.trace_adjoint(mesh, w["Minv"], w["light"], rays, n, d_hits, d_target, d_gv, d_gm, d_loss)
        outs.append((d_gv.download(np.float32).reshape(-1, 3), d_gm.download(np.float32, 16)))
    assert outs[0][0].tobytes() == outs[1][0].tobytes() and outs[0][1].tobytes() == outs[1][1].tobytes()
    hits = d_hits.download(B.HIT_DTYPE, n)
    per_face = np.bincount(hits["prim"][hits["prim"] != B.MISS], minlength=mesh.bvh.len_primitives())
    assert per_face.max() > 64 * 40  # more than 32 waves of records: the workgroup sort
    vid, _, _, _ = vid_of(mesh)
    _, N = mesh.download()
    target = d_target.download(np.float32, 2 * n).reshape(-1, 2)
    _, ref_gv, ref_gm, _ = orc.diff_grad64(w["V"], vid, N, w["Minv"], w["light"],
                                           rays.download(B.RAY_DTYPE, n).view(orc.RAY_DTYPE), hits["prim"], target,
                                           1.0, 1.0)
    per_vertex, normwise, per_entry = gradient_errors(outs[0][0], outs[0][1], ref_gv, ref_gm)
    assert per_vertex.max() <= 1e-4 and normwise <= 1e-6 and per_entry.max() <= 1e-4, \
        (per_vertex.max(), normwise, per_entry.max())


def grazing_cosine(scene, mesh, prims):
    """|cos| of the angle between each ray (in model space) and the normal of
    the triangle it hit (float64; 0 for misses)."""
    _, _, tris = mesh.bvh.download()
    rays = scene["rays"]
    A = scene["w"]["Minv"].astype(np.float64).reshape(4, 4).T
    d = rays["d"].astype(np.float64) @ A[:3, :3].T
    hit = prims != B.MISS
    t = tris[prims[hit]].astype(np.float64)
    n = np.cross(t[:, 3:6] - t[:, 0:3], t[:, 6:9] - t[:, 0:3])
    c = np.zeros(prims.shape[0])
    c[hit] = np.abs((n * d[hit]).sum(1)) / np.linalg.norm(n, axis=1) / np.linalg.norm(d[hit], axis=1)
    return c


def test_adjoint_matches_central_differences_of_gpu_forward(ctx, scene):
    """h = 1e-3 on vertex coordinates and entries of M^-1, normals held fixed.
    Only non-silhouette pixels enter: the ray meets its triangle at more than
    ~75 degrees from grazing (|cos| > 0.25; near a silhouette a step of 1e-3
    changes the grazing angle by O(1) of itself, so differences there say
    nothing about the derivative), and the hit triangle is the same at x - h,
    x and x + h.  The adjoint runs on that same pixel set.  The forward is
    binary32, so the differences carry rounding noise of about 1e-3
    relative.  Probed: the three vertices with the largest gradients and nine
    seeded vertices among those with at least 1% of the largest, each
    compared on its own scale; entries of M^-1 to 1e-2 of the largest."""
    mesh, w = scene["mesh"], scene["w"]
    V0 = w["V"].copy()
    target = scene["target"].astype(np.float64)
    d_hits, _ = forward(ctx, scene)
    hits0 = d_hits.download(B.HIT_DTYPE, NRAY)
    base_prim = hits0["prim"]
    calm = (base_prim != B.MISS) & (target[:, 0] < 3e38) & (grazing_cosine(scene, mesh, base_prim) > 0.25)
    sub = np.nonzero(calm)[0]
    d_gv, _, _ = adjoint(ctx, scene, ctx.to_device(scene["rays"][sub]), sub.size, ctx.to_device(hits0[sub]),
                         ctx.to_device(scene["target"][sub]))
    g_calm = d_gv.download(np.float32).reshape(-1, 3)
    # the three largest gradients, and nine seeded vertices among those whose
    # gradient is at least 1% of the largest
    mag = np.linalg.norm(g_calm, axis=1)
    verts = list(np.argsort(-mag)[:3])
    pool = np.setdiff1d(np.nonzero(mag >= 1e-2 * mag.max())[0], verts)
    verts += list(np.random.default_rng(0xFD).choice(pool, 9, replace=False))
    h = np.float32(1e-3)

    def run(Vx=None, minv=None):
        if Vx is not None:
            mesh.set_vertices(Vx, normals=False)
        dh, di = forward(ctx, scene, minv=minv)
        return dh.download(B.HIT_DTYPE, NRAY)["prim"], di.download(np.float32, 2 * NRAY).reshape(-1, 2)

    def check(prims, imgs, step, grad_of):
        stable = calm.copy()
        for p in prims:
            stable &= p == base_prim
        idx = np.nonzero(stable)[0]
        lp, lm = [((im[idx].astype(np.float64) - target[idx]) ** 2).sum() for im in imgs]
        fd = (lp - lm) / step
        gv, gm, _ = adjoint(ctx, scene, ctx.to_device(scene["rays"][idx]), idx.size, ctx.to_device(hits0[idx]),
                            ctx.to_device(scene["target"][idx]))
        return fd, grad_of(gv.download(np.float32).reshape(-1, 3), gm.download(np.float32, 16)), idx.size

    results = []
    try:
        for v in verts:
            for c in range(3):
                Vp, Vm = V0.copy(), V0.copy()
                Vp[v, c] += h
                Vm[v, c] -= h
                pp, ip = run(Vp)
                pm, im = run(Vm)
                step = float(Vp[v, c]) - float(Vm[v, c])
                results.append(check([pp, pm], [ip, im], step, lambda gv, gm, v=v, c=c: gv[v, c]))
    finally:
        mesh.set_vertices(V0, normals=False)
    minv0 = w["Minv"].astype(np.float32)
    for e in (12, 13, 14, 0, 5):
        mp, mm = minv0.copy(), minv0.copy()
        mp[e] += h
        mm[e] -= h
        pp, ip = run(minv=mp)
        pm, im = run(minv=mm)
        step = float(mp[e]) - float(mm[e])
        results.append(check([pp, pm], [ip, im], step, lambda gv, gm, e=e: gm[e]))
    fd = np.array([r[0] for r in results])
    an = np.array([r[1] for r in results])
    nv = 3 * len(verts)
    vert_fd, vert_an = fd[:nv].reshape(-1, 3), an[:nv].reshape(-1, 3)
    # each probed vertex on its own scale: the difference vector against the
    # adjoint vector, relative to the adjoint vector's length
    rel = np.linalg.norm(vert_fd - vert_an, axis=1) / np.linalg.norm(vert_an, axis=1)
    print("finite differences vs adjoint, per probed vertex:", np.array2string(rel, precision=3))
    assert rel.max() <= 2e-2, rel
    assert np.abs(vert_fd - vert_an).max() <= 1e-2 * np.abs(vert_an).max(), (vert_fd, vert_an)
    m_fd, m_an = fd[nv:], an[nv:]
    assert np.abs(m_fd - m_an).max() <= 1e-2 * np.abs(m_an).max(), (m_fd, m_an)
    assert min(r[2] for r in results) > 100_000


def test_sgd_update_gathers_refits_and_renormalises(ctx, orc, scene):
    w = scene["w"]
    V0 = w["V"].copy()
    mesh = scene["mesh"]
    nodes0, used, _ = mesh.bvh.download()
    d_hits, _ = forward(ctx, scene)
    d_gv, _, _ = adjoint(ctx, scene, scene["d_rays"], NRAY, d_hits, scene["d_target"])
    g = d_gv.download(np.float32).reshape(-1, 3)
    lr = np.float32(2e-3)
    try:
        mesh.sgd(d_gv, float(lr))
        mesh.update(normals=True)
        V1, N1 = mesh.download()
        assert V1.tobytes() == (V0 - lr * g).astype(np.float32).tobytes()
        assert N1.tobytes() == orc.mesh_normals(V1, w["F"]).tobytes()
        vid, nodes1, used1, tris1 = vid_of(mesh)
        assert tris1.tobytes() == np.ascontiguousarray(V1[vid.reshape(-1, 3)].reshape(-1, 9)).tobytes()
        ref = orc.bvh_refit(nodes0.copy(), used, tris1)
        assert nodes1.tobytes() == ref.tobytes()
    finally:
        mesh.set_vertices(V0, normals=True)


def test_gradient_descent_lowers_the_loss(ctx, scene):
    """20 steps of forward, adjoint, clipped SGD on the vertices, refit: the
    loss falls by at least a third.  A fresh mesh keeps the shared one
    intact."""
    w = scene["w"]
    mesh, _ = make_mesh(ctx, w["V"], w["F"])
    d_hits = ctx.alloc(NRAY * B.HIT_DTYPE.itemsize)
    d_img = ctx.alloc(NRAY * 8)
    d_gv = ctx.alloc(mesh.n_vertices * 12)
    d_gm = ctx.alloc(64)
    d_loss = ctx.alloc(8)
    losses = []
    for _ in range(20):
        B.trace_forward(mesh, w["Minv"], w["light"], scene["d_rays"], NRAY, d_hits, d_img)
        B.trace_adjoint(mesh, w["Minv"], w["light"], scene["d_rays"], NRAY, d_hits, scene["d_target"], d_gv, d_gm,
                        d_loss)
        losses.append(float(d_loss.download(np.float64, 1)[0]))
        mesh.sgd(d_gv, scenes.DIFF_LR, scenes.DIFF_MAX_STEP)
        mesh.update(normals=True)
    assert losses[-1] < losses[0] * 2.0 / 3.0, losses
    mesh.close()
