"""GPU tests of the learned warm start: the MFMA network against the FP64
oracle, the FP32 training gradient against the oracle's, training determinism,
that logging and residual measurement leave the simulation bit-identical, and
that a trained network reaches the plain 8-pass contact residual with half the
solver passes."""
import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import _lib
from mechanic_amd.physics import DeviceError, GpuPhysics, GpuPhysicsConfig
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from mechanic_amd import learned
    if mh.device_count() == 0:
        pytest.fail("no HIP device visible")
    return learned


def _inputs(net, rows, rng):
    mean = net[mo.OFF_IN_MEAN:mo.OFF_IN_SCALE]
    scale = net[mo.OFF_IN_SCALE:mo.OFF_OUT_MEAN]
    return (rng.normal(0, 1, (rows, mo.N_IN)) / scale + mean).astype(np.float32)


@pytest.mark.parametrize("rows", [1, 15, 16, 17, 1003, 40000])
def test_mfma_forward_matches_bf16_oracle(L, rows):
    """BF16 operands, FP32 accumulation: equal to the oracle that rounds at the
    same points up to FP32 summation order and the hardware exp/reciprocal in
    GELU. Those can move a rare hidden value across a BF16 rounding boundary
    (one BF16 step, 2^-8 relative), which bounds the worst row; a wrong operand
    layout would instead be wrong by O(1) in most rows."""
    rng = np.random.default_rng(rows)
    net = mo.random_net(rng).astype(np.float32)
    X = _inputs(net, rows, rng)
    w = L.WarmStartNet(0)
    w.set_net(net)
    y = w.predict(X).astype(np.float64)
    scale = net[mo.OFF_OUT_SCALE:mo.N_NET].astype(np.float64)
    eb = np.abs(y - mo.forward(net, X, bf16_mode=True)) / scale
    assert eb.max() < 2e-2, eb.max()
    assert np.quantile(eb, 0.999) < 1e-4, np.quantile(eb, 0.999)
    assert np.median(eb) < 1e-6, np.median(eb)
    # Against exact FP64: BF16 operands leave a few per cent at most.
    ee = np.abs(y - mo.forward(net, X)) / scale
    assert ee.max() < 0.1 and np.median(ee) < 0.01, (ee.max(), np.median(ee))


def test_mfma_single_path_network(L):
    """The hand-built single-path network: every weight and the chosen inputs
    are exact in BF16, so the MFMA result equals the closed form up to the BF16
    rounding of the two hidden activations."""
    net = np.zeros(mo.N_NET, np.float32)
    net[mo.OFF_IN_SCALE:mo.OFF_OUT_MEAN] = 1.0
    net[mo.OFF_OUT_SCALE:mo.N_NET] = 1.0
    net[mo.OFF_W1 + 3 * mo.N_H + 5] = 1.5
    net[mo.OFF_W2 + 5 * mo.N_H + 41] = -0.75
    net[mo.OFF_W3 + 41 * mo.N_OUT + 1] = 2.0
    net[mo.OFF_B3:mo.N_PARAMS] = [0.25, -0.5, 0.125]
    X = np.zeros((33, mo.N_IN), np.float32)
    X[:, 3] = np.linspace(-2, 2, 33)
    w = L.WarmStartNet(0)
    w.set_net(net)
    y = w.predict(X).astype(np.float64)
    expect = mo.forward(net, X, bf16_mode=True)
    np.testing.assert_allclose(y, expect, rtol=0, atol=1e-5)
    np.testing.assert_allclose(y[:, 0], 0.25)
    np.testing.assert_allclose(y[:, 2], 0.125)


def test_dataset_normalisation(L):
    rng = np.random.default_rng(4)
    X = rng.normal(2.0, 3.0, (5000, mo.N_IN)).astype(np.float32)
    X[:, 9] = 0.5
    T = rng.normal(-1.0, 0.25, (5000, mo.N_OUT)).astype(np.float32)
    w = L.WarmStartNet(0)
    w.set_dataset(X, T)
    got = w.get_net().astype(np.float64)
    s = mo.normalisation(X, T)
    np.testing.assert_allclose(got[mo.OFF_IN_MEAN:mo.OFF_IN_SCALE], s["in_mean"], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got[mo.OFF_IN_SCALE:mo.OFF_OUT_MEAN], s["in_scale"], rtol=1e-6)
    np.testing.assert_allclose(got[mo.OFF_OUT_MEAN:mo.OFF_OUT_SCALE], s["out_mean"], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got[mo.OFF_OUT_SCALE:mo.N_NET], s["out_scale"], rtol=1e-6)


@pytest.mark.parametrize("rows", [1, 31, 32, 33, 2500])
def test_gradient_matches_fp64_oracle(L, rows):
    """FP32 forward/backward on the GPU against the FP64 oracle, whose gradient
    is pinned by finite differences on the CPU."""
    rng = np.random.default_rng(100 + rows)
    net = mo.random_net(rng).astype(np.float32)
    X = _inputs(net, rows, rng)
    T = rng.normal(0, 1, (rows, mo.N_OUT)).astype(np.float32)
    w = L.WarmStartNet(0)
    w.set_dataset(X, T)
    w.set_net(net)  # replace the dataset normalisation with the test network's
    g, loss = w.gradient()
    ref_loss, ref = mo.loss_and_grad(net, X, T)
    assert abs(loss - ref_loss) <= 1e-5 * max(1.0, ref_loss)
    assert np.linalg.norm(g - ref) <= 1e-5 * np.linalg.norm(ref)
    assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()


def test_training_fits_a_teacher_and_is_deterministic(L):
    rng = np.random.default_rng(8)
    teacher = mo.random_net(np.random.default_rng(9))
    X = rng.normal(0, 1, (20000, mo.N_IN)).astype(np.float32)
    T = mo.forward(teacher, X).astype(np.float32)
    a = L.WarmStartNet(5)
    la = a.fit(X, T, epochs=8, batch=1024, learning_rate=2e-3)
    b = L.WarmStartNet(5)
    lb = b.fit(X, T, epochs=8, batch=1024, learning_rate=2e-3)
    assert np.array_equal(la, lb)
    assert np.array_equal(a.get_net(), b.get_net())
    assert la[-1] < 0.25 * la[0], la
    # The loss the trainer reports is the oracle's loss of the parameters it used.
    _, final = a.gradient()
    assert abs(final - mo.loss_only(a.get_net().astype(np.float64), X, T)) < 1e-4 * max(1.0, final)


def _state_after(scene, cfg, ticks, net=None):
    with GpuPhysics.new_with_config(scene, cfg) as g:
        if net is not None:
            g.set_mlp(net)
        for t in range(1, ticks + 1):
            g.dispatch_tick(t)
        assert g.read_last_tick().error_flags == 0
        return g.read_bodies()


def test_logging_and_residual_do_not_change_the_simulation(L):
    scene = mh.cuboid_pile(8, 6, 8, seed=3)
    base = GpuPhysicsConfig(friction=0.6, solver_iterations=8)
    p0, v0 = _state_after(scene, base, 60)
    for cfg in (GpuPhysicsConfig(friction=0.6, solver_iterations=8, warm_start=_lib.WARM_START_LOG),
                GpuPhysicsConfig(friction=0.6, solver_iterations=8, residual=True)):
        p1, v1 = _state_after(scene, cfg, 60)
        assert np.array_equal(p0.view(np.uint32), p1.view(np.uint32))
        assert np.array_equal(v0.view(np.uint32), v1.view(np.uint32))


def test_training_rows_match_their_definition(L):
    """Features of the logged tick agree with the state they describe: the
    ground flag with the contact rows, the cache-hit flag is 0/1, masses and
    inverse inertias are those of the scene, targets are finite and the
    normal impulse is non-negative."""
    scene = mh.cuboid_pile(6, 4, 6, seed=2)
    cfg = GpuPhysicsConfig(friction=0.6, solver_iterations=8, warm_start=_lib.WARM_START_LOG)
    with GpuPhysics.new_with_config(scene, cfg) as g:
        for t in range(1, 81):
            g.dispatch_tick(t)
        F, T = g.read_training_rows()
        rb = g.read_last_tick()
    assert F.shape[0] == rb.active_contact_count > 0
    assert np.all(np.isfinite(F)) and np.all(np.isfinite(T))
    assert set(np.unique(F[:, 14])) <= {0.0, 1.0} and set(np.unique(F[:, 15])) <= {0.0, 1.0}
    inv_mass = float(scene.inverse_masses[0])
    assert np.all(F[:, 4] == inv_mass)
    assert np.all(F[F[:, 15] == 0.0, 5] == inv_mass) and np.all(F[F[:, 15] == 1.0, 5] == 0.0)
    assert np.all(T[:, 0] >= 0.0)
    # After settling most contacts find their cached manifold.
    assert F[:, 14].mean() > 0.5


def test_learned_mode_requires_a_network(L):
    scene = mh.cuboid_pile(3, 2, 3, seed=0)
    with GpuPhysics.new_with_config(scene, GpuPhysicsConfig(warm_start=_lib.WARM_START_LEARNED)) as g:
        with pytest.raises(DeviceError):
            g.dispatch_tick(1)


def test_learned_warm_start_halves_solver_passes(L):
    """Trained on one pile (seed 7), evaluated on another (seed 1): with the
    learned warm start, 4 solver passes leave a lower mean contact residual
    than the reference's cache warm start with 8, and the run stays finite."""
    friction = 0.6
    train = mh.cuboid_pile(20, 12, 20, seed=7)
    F, T = L.collect_rows(train, GpuPhysicsConfig(friction=friction, solver_iterations=32), 200,
                          range(20, 201, 4))
    net = L.WarmStartNet(3)
    losses = net.fit(F, T, epochs=20, batch=4096, learning_rate=1e-3)
    assert losses[-1] < losses[0]
    scene = mh.cuboid_pile(20, 12, 20, seed=1)
    plain = L.residual_trace(scene, GpuPhysicsConfig(friction=friction, solver_iterations=8), 200)
    learned = L.residual_trace(scene, GpuPhysicsConfig(friction=friction, solver_iterations=4), 200, net=net)
    print(f"mean residual: plain 8 passes {plain.mean():.4f}, learned 4 passes {learned.mean():.4f}")
    assert np.all(np.isfinite(learned))
    assert learned.mean() < plain.mean()


def _trained(L, seed=7):
    train = mh.cuboid_pile(12, 8, 12, seed=seed)
    F, T = L.collect_rows(train, GpuPhysicsConfig(friction=0.6, solver_iterations=32), 120, range(20, 121, 10))
    net = L.WarmStartNet(3)
    net.fit(F, T, epochs=10, batch=4096, learning_rate=1e-3)
    return net


@pytest.mark.parametrize("solve_kind", [0, 1])
def test_non_finite_predictions_fall_back_to_the_cache_warm_start(L, solve_kind):
    """A network whose outputs are not finite (NaN weights): every contact takes
    the cache warm start and the tick runs its fallback passes, so 4 learned
    passes plus 4 fallback passes give a plain 8-pass run bit for bit, and the
    fallback counts say so."""
    scene = mh.cuboid_pile(8, 6, 8, seed=3)
    net = L.WarmStartNet(1)
    bad = net.get_net()
    bad[:net.param_count] = np.nan
    net.set_net(bad)
    plain = GpuPhysicsConfig(friction=0.6, solver_iterations=8)
    learned = GpuPhysicsConfig(friction=0.6, solver_iterations=4, warm_start=_lib.WARM_START_LEARNED,
                               learned_residual_limit=1e9, learned_fallback_passes=4)
    with GpuPhysics.new_with_config(scene, plain) as g:
        g.set_tuning(solve_kind=solve_kind)
        for t in range(1, 41):
            g.dispatch_tick(t)
        p0, v0 = g.read_bodies()
    with GpuPhysics.new_with_config(scene, learned) as g:
        g.set_tuning(solve_kind=solve_kind)
        g.set_mlp(net)
        for t in range(1, 41):
            g.dispatch_tick(t)
        p1, v1 = g.read_bodies()
        stats = g.read_learned_stats()
        assert g.read_last_tick().error_flags == 0
    assert np.array_equal(p0.view(np.uint32), p1.view(np.uint32))
    assert np.array_equal(v0.view(np.uint32), v1.view(np.uint32))
    assert stats["fallback_ticks"] == 40 and stats["fallback_this_tick"] == 1
    assert stats["nonfinite_this_tick"] > 0 and stats["nonfinite_total"] >= stats["nonfinite_this_tick"]


def test_residual_limit_decides_the_fallback(L):
    """With a limit no tick reaches the extra passes never run (the run equals
    one without the check); with a zero-width margin below every tick's
    residual they run every tick, and the residual the tick reports drops."""
    net = _trained(L)
    scene = mh.cuboid_pile(12, 8, 12, seed=1)
    base = dict(friction=0.6, solver_iterations=4, warm_start=_lib.WARM_START_LEARNED, residual=True)
    runs = {}
    for name, extra in (("none", {}), ("never", dict(learned_residual_limit=1e9)),
                        ("always", dict(learned_residual_limit=1e-12, learned_fallback_passes=4))):
        with GpuPhysics.new_with_config(scene, GpuPhysicsConfig(**base, **extra)) as g:
            g.set_mlp(net)
            rms = []
            passes = []
            for t in range(1, 61):
                g.dispatch_tick(t)
                r = g.read_residual()
                rms.append(r["rms"])
                passes.append(r["fallback_passes"])
            runs[name] = (g.read_bodies(), np.asarray(rms), passes, g.read_learned_stats())
    (p_none, v_none), rms_none, _, _ = runs["none"]
    (p_never, v_never), rms_never, passes_never, st_never = runs["never"]
    (_, _), rms_always, passes_always, st_always = runs["always"]
    assert np.array_equal(p_none.view(np.uint32), p_never.view(np.uint32))
    assert np.array_equal(v_none.view(np.uint32), v_never.view(np.uint32))
    assert st_never["fallback_ticks"] == 0 and set(passes_never) == {0}
    ticks_with_contacts = int((rms_none > 0).sum())
    assert st_always["fallback_ticks"] >= ticks_with_contacts and set(passes_always) <= {0, 4}
    print(f"mean residual: learned 4 {rms_none.mean():.4f}, with 4 fallback passes every tick {rms_always.mean():.4f}")
    assert rms_always.mean() < rms_none.mean()
