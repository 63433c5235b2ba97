"""GPU rigid-body physics (physics/, Scene::run) against the oracle: body
records bit for bit after every step, object transforms and the traced
scene bit for bit after Scene::run."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import physics as P
from bvhtracer_amd import scenes
from helpers import assert_hits_equal

pytestmark = pytest.mark.gpu
f32 = np.float32


def random_world(ctx, n, n_springs, seed):
    rng = np.random.default_rng(seed)
    world = P.World(ctx)
    inst = []
    for i in range(n):
        b = P.RigidBody()
        kind = i % 5
        if kind == 1:
            b.set_mass(rng.uniform(1, 5))
        elif kind == 2:
            b.set_inverse_mass(0.0)  # infinite mass: Gravity applies f32::MAX
        elif kind == 3:
            b.set_mass(rng.uniform(0.5, 2))
            b.set_inverse_inertia_tensor(np.diag(rng.uniform(0.2, 3, 3)).astype(f32))
        b.set_drag(rng.choice([1.0, rng.uniform(0.3, 1.0)]), rng.choice([1.0, rng.uniform(0.3, 1.0)]))
        b.set_position(rng.uniform(-5, 5, 3))
        b.set_orientation(rng.normal(size=4))
        b.set_velocity(rng.uniform(-1, 1, 3))
        b.set_rotation(rng.uniform(-2, 2, 3))
        b.set_acceleration(rng.uniform(-0.5, 0.5, 3) if i % 7 == 0 else (0, 0, 0))
        if i % 3 == 0:
            b.set_awake(True)
        if i % 4 == 0:
            b.set_can_sleep(False)
        inst.append(world.register_body(b))
    for i in range(0, n, 2):
        world.register_force_generator(inst[i], P.Gravity(rng.uniform(-10, 10, 3)))
    for i in range(0, n, 9):
        world.register_force_generator(inst[i], P.NullForce())
    finite = [i for i in range(n) if i % 5 in (1, 3)]  # springs act on bodies with a set mass
    for _ in range(n_springs):
        a, c = rng.choice(finite), rng.choice([i for i in range(n) if i % 5 != 2])
        world.register_force_generator(inst[a], P.Spring(inst[c], rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3),
                                                         rng.uniform(0.05, 0.5), rng.uniform(0, 2)))
    forces = np.concatenate(world._forces)
    return world, inst, forces


def canonical(bodies):
    """Body records with every NaN replaced by one pattern: the sign and
    payload of a NaN differ between the host CPU and the GPU and carry no
    meaning."""
    u = bodies.view(np.uint32).copy()
    f = bodies.view(np.float32)
    u[np.isnan(f)] = 0x7FC00000
    return u.tobytes()


def test_world_steps_bit_identical(ctx, orc):
    world, _, forces = random_world(ctx, 300, 150, seed=1)
    ref = world.bodies().view(orc.BODY_DTYPE).copy()
    world.start_frame()
    orc.world_start_frame(ref)
    assert canonical(world.bodies()) == canonical(ref), "start_frame"
    durations = [1 / 60, 1 / 30, 0.1, 0.25, 1 / 60, 0.5, 1.0, 0.0, 1 / 120] * 4
    for k, dt in enumerate(durations):
        world.run_physics(dt)
        orc.world_run_physics(ref, forces, dt)
        got = world.bodies()
        if canonical(got) != canonical(ref):
            bad = np.nonzero(got.view(np.uint8).reshape(len(got), -1).any(1) !=
                             ref.view(np.uint8).reshape(len(ref), -1).any(1))[0]
            diff = [i for i in range(len(got)) if got[i].tobytes() != ref[i].tobytes()]
            raise AssertionError(f"step {k} (dt {dt}): bodies {diff[:10]} differ {bad[:3]}")
    end = world.bodies()
    awake = end["is_awake"]
    assert 0 < awake.sum() < len(awake)  # both branches of the sleep rule ran
    # infinite-mass bodies under gravity go to NaN as in the reference (g * f32::MAX, then inf * 0)
    sane = np.array([i % 5 != 2 for i in range(len(end))])
    assert np.isfinite(end["position"][sane]).all() and np.isfinite(end["orientation"][sane]).all()
    assert np.isnan(end["velocity"][~sane]).any()


def test_body_handles_read_and_write_device_state(ctx, orc):
    world = P.World(ctx)
    b = P.RigidBody()
    b.set_mass(1.0)
    b.set_can_sleep(False)
    h = world.register_body(b)
    world.register_force_generator(h, P.Gravity((0, -8, 0)))
    for _ in range(3):
        world.run_physics(0.5)
    assert h.borrow().get_velocity().tolist() == [0.0, -12.0, 0.0]
    assert h.borrow().get_position().tolist() == [0.0, -12.0, 0.0]
    with h.borrow_mut() as body:
        body.set_velocity((0, 0, 0))
        body.set_position((1, 2, 3))
    world.run_physics(0.5)
    assert h.borrow().get_velocity().tolist() == [0.0, -4.0, 0.0]
    assert h.borrow().get_position().tolist() == [1.0, 0.0, 3.0]


def unit_cube():
    """Twelve triangles of [0, 1]^3."""
    v = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], f32)
    faces = [(0, 1, 3), (0, 3, 2), (4, 6, 7), (4, 7, 5), (0, 4, 5), (0, 5, 1), (2, 3, 7), (2, 7, 6),
             (0, 2, 6), (0, 6, 4), (1, 5, 7), (1, 7, 3)]
    return np.stack([v[list(f)].reshape(9) for f in faces]).astype(f32)


def test_scene_run_matches_oracle(ctx, orc):
    """examples/cube.rs and two_armadillos.rs set-ups: Scene::run moves the
    objects by their bodies; transforms, bounds, TLAS and hits match the
    oracle after every frame."""
    cube = B.ModelBuilder(ctx=ctx).with_mesh(unit_cube()).build()
    sphere = B.ModelBuilder(ctx=ctx).with_mesh(scenes.icosphere(8, radius=0.8)).with_build(B.BUILD_SAH, 2).build()
    world = P.World(ctx)
    b1 = P.RigidBody()
    b1.set_position((0, 0, 0))
    b1.set_orientation((0, 0, 0, 1))  # Quaternion::unit_z
    b1.set_rotation((0, 0, 2))
    b1.set_awake(True)
    b1.set_angular_drag(1.0)
    h1 = world.register_body(b1)
    b2 = P.RigidBody()
    b2.set_position((1.3, 0, 0))
    b2.set_orientation((1, 0, 0, 0))
    b2.set_rotation((0, 0.02, 0))
    b2.set_angular_drag(1.0)
    b2.set_can_sleep(False)
    b2.set_awake(True)
    h2 = world.register_body(b2)
    b3 = P.RigidBody()
    b3.set_mass(2.0)
    b3.set_position((-2.2, 0.5, 0))
    b3.set_orientation((1, 0, 0, 0))
    b3.set_can_sleep(False)
    h3 = world.register_body(b3)
    world.register_force_generator(h3, P.Gravity((0, -0.5, 0)))
    world.register_force_generator(h3, P.Spring(h2, (0.2, 0, 0), (0, 0.3, 0), 3.0, 0.5))
    scale_translate = np.diag([2, 2, 2, 1]).astype(f32)
    scale_translate[3, :3] = (-1, -1, -1)
    static = np.diag([0.5, 0.5, 0.5, 1]).astype(f32)
    static[3, :3] = (0.8, -2.0, 0.0)
    objects = [B.SceneObjectBuilder(cube, h1).with_transform(scale_translate).build(),
               B.SceneObjectBuilder(sphere, h2).build(),
               B.SceneObjectBuilder(sphere, h3).with_transform(np.diag([1, 1, 1, 1]).astype(f32)).build(),
               B.SceneObjectBuilder(cube).with_transform(static).build()]
    scene = B.SceneBuilder(ctx=ctx).with_physics(world).with_objects(objects).build()
    ref = world.bodies().view(orc.BODY_DTYPE).copy()
    forces = np.concatenate(world._forces)
    object_body = np.array([0, 1, 2, 0xFFFFFFFF], np.uint32)
    init = np.stack([o.transform_init for o in objects])
    blas_arrays = [m.bvh.download() for m in scene.models]
    blases = [(nodes, tris) for nodes, _, tris in blas_arrays]
    cam = scenes.camera(eye=(0, 0.5, -6), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=50)
    rays = orc.raygen(cam, 160, 160)
    prays = np.zeros(rays.shape[0], B.RAY_DTYPE)
    prays.view(np.uint8)[:] = rays.view(np.uint8)
    static_m = objects[3].m.copy()
    for frame in range(12):
        elapsed = 1 / 60 if frame % 3 else 0.05
        scene.run(elapsed)
        orc.world_run_physics(ref, forces, elapsed)
        m, minv = orc.objects_from_bodies(ref, object_body, init)
        m[3], minv[3] = static_m, objects[3].minv
        for i in range(4):
            o = scene.get_unchecked(i)
            assert o.m.tobytes() == m[i].tobytes() and o.minv.tobytes() == minv[i].tobytes(), f"frame {frame} obj {i}"
        assert world.bodies().tobytes() == ref.tobytes(), f"frame {frame} bodies"
        bounds = np.stack([np.concatenate(orc.instance_bounds(m[i], blases[objects[i].model][0][0]["bmin"],
                                                              blases[objects[i].model][0][0]["bmax"]))
                           for i in range(4)])
        tnodes, used = orc.tlas_build(bounds)
        g_nodes, g_used = scene.tlas.download()
        assert g_used == used and g_nodes.tobytes() == tnodes.tobytes(), f"frame {frame} TLAS"
        inst = np.zeros(4, orc.INSTANCE_DTYPE)
        inst["minv"], inst["blas"] = minv, [o.model for o in objects]
        want = orc.tlas_trace(tnodes, inst, blases, rays, nthreads=8)
        got = scene.intersect(prays)
        assert_hits_equal(got, want, f"frame {frame}")
    assert len(np.unique(got["inst"][got["prim"] != B.MISS])) >= 3


def test_scene_run_captured_in_a_graph(ctx, orc):
    """Scene::run_physics + asynchronous TLAS rebuild + trace replayed as a
    hipGraph gives the same bodies, TLAS and hits as eager calls."""
    cube = B.ModelBuilder(ctx=ctx).with_mesh(unit_cube()).build()

    def make():
        w = P.World(ctx)
        hs = []
        for k in range(6):
            b = P.RigidBody()
            b.set_position((k - 2.5, 0, 0))
            b.set_orientation((1, 0, 0, 0))
            b.set_rotation((0.3 * k, 0.1, 0.7))
            b.set_can_sleep(False)
            b.set_awake(True)
            hs.append(w.register_body(b))
        objs = [B.SceneObjectBuilder(cube, h).build() for h in hs]
        return w, B.SceneBuilder(ctx=ctx).with_physics(w).with_objects(objs).build()

    w_e, s_e = make()
    w_g, s_g = make()
    cam = scenes.camera(eye=(0, 0, -8), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=60)
    n = 96 * 96
    d_rays = B.raygen(ctx, cam, 96, 96, B.ORDER_ROWS)
    d_he, d_hg = ctx.alloc(n * B.HIT_DTYPE.itemsize), ctx.alloc(n * B.HIT_DTYPE.itemsize)
    s_g.run_physics(0.0)  # bind and warm up outside the capture
    s_e.run_physics(0.0)
    s_e.tlas.rebuild_async()
    s_g.tlas.rebuild_async()
    s_g.tlas.intersect_device(d_rays, n, d_hg)  # first trace outside the capture
    s_e.tlas.intersect_device(d_rays, n, d_he)
    g = B.Graph.capture(ctx, lambda: (s_g.run_physics(1 / 30), s_g.tlas.rebuild_async(),
                                      s_g.tlas.intersect_device(d_rays, n, d_hg)))
    for frame in range(5):
        s_e.run_physics(1 / 30)
        s_e.tlas.rebuild_async()
        s_e.tlas.intersect_device(d_rays, n, d_he)
        g.launch()
        ctx.sync()
        assert w_g.bodies().tobytes() == w_e.bodies().tobytes(), f"replay {frame} bodies"
        he, hg = d_he.download(B.HIT_DTYPE, n), d_hg.download(B.HIT_DTYPE, n)
        assert hg.tobytes() == he.tobytes(), f"replay {frame} hits"
    assert (he["prim"] != B.MISS).sum() > 100
