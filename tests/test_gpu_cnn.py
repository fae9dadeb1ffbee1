"""GPU tests of the learned preconditioner: BF16 MFMA convolution kernels against the pinned
CPU oracle (forward activations, weight gradients), the training set's pairs against the
matrix (A e = r), training, and flexible CG with the network -- its gain over plain CG on the
same system, and its fall-back to plain FP64 CG when the network is useless.

Tolerances: activations are BF16 (8-bit mantissa); the GPU's FP32 MFMA sums are ordered
differently from the oracle's double sums, so a stored BF16 value may differ by one BF16 ulp
(2^-8 relative) where the sum sits near a rounding boundary; such flips are rare and bounded.
"""
import math

import numpy as np
import pytest

import hpccg_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hpccg_amd.Context(0)
    yield c
    c.close()


def he_weights(seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.standard_normal(oracle.NW1) * math.sqrt(2 / 27),
                           rng.standard_normal(oracle.NW2) * math.sqrt(2 / 432),
                           rng.standard_normal(oracle.NW3) * math.sqrt(1 / 432)]).astype(np.float32)


def bf16_close(a, b, frac=0.01):
    """Equal as BF16 except for a small fraction of one-ulp differences."""
    diff = np.abs(a - b)
    tol = np.abs(b) * 2.0 ** -7 + 1e-30
    assert np.all(diff <= tol + 1e-6 * np.abs(b).max()), float((diff - tol).max())
    assert np.mean(diff > 0) <= frac, float(np.mean(diff > 0))


@pytest.mark.parametrize("dims,batch", [((8, 8, 4), 1), ((13, 9, 6), 1), ((10, 7, 5), 2), ((16, 16, 16), 1)])
def test_forward_matches_oracle(ctx, dims, batch):
    w = he_weights(1)
    ctx.cnn_set_weights(w)
    assert np.array_equal(ctx.cnn_get_weights(), w)
    n = int(np.prod(dims)) * batch
    u = np.random.default_rng(2).standard_normal(n).astype(np.float32)
    z, h1, h2 = ctx.cnn_forward(u, dims, batch, want_act=True)
    zo, h1o, h2o = oracle.cnn_forward(w, u, dims, batch, quant=True)
    bf16_close(h1, h1o)
    bf16_close(h2, h2o, frac=0.02)
    assert np.linalg.norm(z - zo) <= 2e-2 * np.linalg.norm(zo)


def test_mfma_operand_layout_single_tap():
    """A weight set with one nonzero tap per layer turns the network into shifts and channel
    selections, so each MFMA operand-layout error shows up as a wrong voxel or channel."""
    c = hpccg_amd.Context(0)
    try:
        dims = (8, 8, 4)
        w = np.zeros(oracle.NW1 + oracle.NW2 + oracle.NW3, np.float32)
        t_center = 13
        w[3 * 27 + 22] = 1.0  # W1[co=3][t=22]: h1[3](v) = relu(u(v + (1,1,1)... ))
        w[oracle.NW1 + (5 * 16 + 3) * 27 + 4] = 1.0  # W2[co=5][ci=3][t=4]
        w[oracle.NW1 + oracle.NW2 + 5 * 27 + t_center] = 1.0  # W3[ci=5][centre]
        c.cnn_set_weights(w)
        u = np.abs(np.random.default_rng(3).standard_normal(256)).astype(np.float32) + 0.5
        z, h1, h2 = c.cnn_forward(u, dims, 1, want_act=True)
        zo, h1o, h2o = oracle.cnn_forward(w, u, dims, 1, quant=True)
        assert np.array_equal(h1, h1o) and np.array_equal(h2, h2o) and np.array_equal(z, zo)
    finally:
        c.close()


@pytest.mark.parametrize("dims", [(16, 16, 16), (37, 21, 19), (5, 40, 3), (48, 32, 70)])
def test_fused_inference_equals_layerwise_forward(ctx, dims):
    """The fused three-layer inference kernel (activations kept in LDS while a workgroup marches
    along z) uses the per-layer kernels' MFMA operands and K order, so its output is bitwise
    theirs -- on grids that are not multiples of its 16 x 16 footprint, with few z-planes, and
    with more z-chunks than one; and it honours the input/output scales."""
    w = he_weights(7)
    ctx.cnn_set_weights(w)
    n = int(np.prod(dims))
    u = np.random.default_rng(8).standard_normal(n).astype(np.float32)
    z_layers = ctx.cnn_forward(u, dims, 1)
    ud = hpccg_amd.DeviceArray.from_numpy(ctx, u)
    zd = hpccg_amd.DeviceArray(ctx, n, np.float32)
    ctx.cnn_apply_dev(dims, ud, 1.0, zd)
    assert np.array_equal(zd.numpy(), z_layers)
    # z = s * net(u / s): the network is positively homogeneous, so only BF16 rounding differs
    s = 4.0  # a power of two: the scaled input rounds exactly like the unscaled one
    ctx.cnn_apply_dev(dims, ud, 1.0 / s, zd)
    assert np.array_equal(zd.numpy(), z_layers)
    zo, _, _ = oracle.cnn_forward(w, u, dims, 1, quant=True)
    assert np.linalg.norm(zd.numpy() - zo) <= 2e-2 * np.linalg.norm(zo)


@pytest.mark.parametrize("dims,batch", [((8, 8, 4), 1), ((11, 6, 9), 2)])
def test_backward_matches_oracle(ctx, dims, batch):
    w = he_weights(4)
    ctx.cnn_set_weights(w)
    n = int(np.prod(dims)) * batch
    u = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    dz = np.random.default_rng(6).standard_normal(n).astype(np.float32)
    g = ctx.cnn_backward(u, dz, dims, batch)
    go = oracle.cnn_backward(w, u, dz, dims, batch, quant=True)
    for lo, hi in [(0, oracle.NW1), (oracle.NW1, oracle.NW1 + oracle.NW2), (oracle.NW1 + oracle.NW2, len(w))]:
        err = np.linalg.norm(g[lo:hi] - go[lo:hi]) / np.linalg.norm(go[lo:hi])
        assert err <= 2e-2, (lo, err)


@pytest.fixture(scope="module")
def trained(ctx):
    """A small training run: pairs from plain CG on 32^3, 16^3 crops."""
    ctx.cnn_init(seed=3)
    ctx.cnn_make_dataset(n=32, npairs=1500, crop=16, seed=21)
    loss, ms = ctx.cnn_train(steps=600, batch=16, lr=2e-3, seed=5)
    return {"loss": loss, "ms": ms, "w": ctx.cnn_get_weights()}


def test_dataset_pairs_satisfy_A_e_eq_r(ctx, trained):
    C = 16
    for i in (0, 37, 700, 1499):
        r, e, m = ctx.cnn_dataset_pair(i, C)
        assert set(m.tolist()) <= {0, 4}
        E = e.reshape(C, C, C).astype(np.float64)
        P = np.zeros((C + 2,) * 3)
        P[1:-1, 1:-1, 1:-1] = E
        box = sum(P[1 + dz:1 + dz + C, 1 + dy:1 + dy + C, 1 + dx:1 + dx + C]
                  for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1))
        Ae = 28 * E - box
        R = r.reshape(C, C, C)
        inner = (slice(1, -1),) * 3  # the crop's own faces see only part of the stencil
        assert np.linalg.norm((Ae - R)[inner]) <= 1e-5 * np.linalg.norm(R[inner]) + 1e-6
        assert abs(np.sqrt(np.mean(r.astype(np.float64) ** 2)) - 1.0) < 1e-5  # scaled by 1/rms(r)


def test_dataset_windows_cover_domain_faces(ctx, trained):
    """Per axis a window lies against the low face, the high face or anywhere (1/3 each), so
    about 1/3 + (1/3)(1/17) of the margins are 0 at 32^3 with 16^3 crops (17 origins)."""
    zeros = [int(x) for i in range(0, 1500, 5) for x in ctx.cnn_dataset_pair(i, 16)[2] if x == 0]
    frac = len(zeros) / (300 * 6)
    assert 0.28 <= frac <= 0.42, frac


def test_training_reduces_loss(trained):
    loss = trained["loss"]
    assert np.all(np.isfinite(loss))
    assert loss[-50:].mean() < 0.5 * loss[:20].mean(), (loss[:20].mean(), loss[-50:].mean())


def _system(n, seed):
    A = oracle.Problem(n, n, n)
    xs = oracle.xstar_sines(n, n, n, seed)
    return A, xs, oracle.spmv(A, xs)


def test_fcg_learned_gain_and_accuracy(ctx):
    """The shipped network (hpccg_amd/data, trained on 64^3 pairs and selected on a 128^3
    validation system) at least halves flexible-CG iterations against plain CG's on a fresh 32^3
    system, and the answer meets the tolerance in the FP64 true residual (checked by the oracle)."""
    from hpccg_amd import learned
    w = learned.load()
    assert w is not None and w.size == oracle.NW1 + oracle.NW2 + oracle.NW3
    ctx.cnn_set_weights(w)
    n = 32
    A, xs, b = _system(n, 777)
    normb = math.sqrt(oracle.ddot1(b))
    ctx.upload(A.rowindex, A.colindex, A.values)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
    res = ctx.fcg_learned_dev((n, n, n), bd, xd, max_iter=500, tolerance=1e-10 * normb)
    x = xd.numpy()
    rel = math.sqrt(oracle.ddot1(b - oracle.spmv(A, x))) / normb
    assert rel <= 1e-10, (rel, res)
    _, nit_plain, _, _, _ = oracle.cg(A, b, max_iter=500, tolerance=1e-10 * normb)
    assert not res["fell_back"], res
    assert res["fcg_iters"] <= 0.5 * nit_plain, (res["fcg_iters"], nit_plain)


@pytest.mark.parametrize("dims", [(36, 20, 28), (17, 45, 9)])
def test_fcg_learned_on_ragged_grids(ctx, dims):
    """Grids whose extents are not multiples of the network's 16 x 16 footprint (partial
    columns in x and y, z-chunks of uneven length): the flexible CG meets 1e-10 in the FP64
    true residual (checked by the oracle) without falling back, in fewer iterations than plain
    CG."""
    from hpccg_amd import learned
    ctx.cnn_set_weights(learned.load())
    A = oracle.Problem(*dims)
    b = oracle.spmv(A, oracle.xstar_sines(*dims, 5))
    normb = math.sqrt(oracle.ddot1(b))
    ctx.upload(A.rowindex, A.colindex, A.values)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
    res = ctx.fcg_learned_dev(dims, bd, xd, max_iter=500, tolerance=1e-10 * normb)
    rel = math.sqrt(oracle.ddot1(b - oracle.spmv(A, xd.numpy()))) / normb
    assert rel <= 1e-10, (rel, res)
    _, nit_plain, _, _, _ = oracle.cg(A, b, max_iter=500, tolerance=1e-10 * normb)
    assert not res["fell_back"] and res["fcg_iters"] < nit_plain, (res, nit_plain)


def test_training_pipeline_selects_a_network_that_beats_plain_cg(ctx):
    """hpccg_amd.learned.train end to end at a reduced size (48^3 pairs, 4 initialisations):
    the selected network needs fewer flexible-CG iterations than plain CG on the validation
    system, without falling back."""
    from hpccg_amd import learned
    recipe = {"n": 48, "npairs": 4000, "crop": 16, "steps": 1500, "batch": 32, "candidates": 4,
              "val_n": 32, "val_seed": 99}
    w, info = learned.train(ctx, recipe, log=lambda *a: None)
    best = min(c["val_iters"] for c in info["candidates"])
    n = recipe["val_n"]
    A = oracle.Problem(n, n, n)
    b = oracle.spmv(A, oracle.xstar_sines(n, n, n, recipe["val_seed"]))
    _, nit_plain, _, _, _ = oracle.cg(A, b, max_iter=500, tolerance=1e-10 * math.sqrt(oracle.ddot1(b)))
    assert best < nit_plain, (info["candidates"], nit_plain)
    assert learned.validation_iters(ctx, w, n, recipe["val_seed"]) == best


@pytest.mark.parametrize("kind", ["zero", "stagnation"])
def test_fcg_falls_back_to_plain_cg(ctx, kind):
    """A useless network hands the solve to plain FP64 CG, which still reaches the tolerance:
    zero weights give z = 0 (breakdown at once); the stagnation case demands a 1e6-fold
    reduction per 10 iterations, which no preconditioner delivers (forced from the host)."""
    w = np.zeros(oracle.NW1 + oracle.NW2 + oracle.NW3, np.float32) if kind == "zero" else he_weights(9)
    ctx.cnn_set_weights(w)
    n = 24
    A, xs, b = _system(n, 5)
    normb = math.sqrt(oracle.ddot1(b))
    ctx.upload(A.rowindex, A.colindex, A.values)
    bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
    xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
    res = ctx.fcg_learned_dev((n, n, n), bd, xd, max_iter=500, tolerance=1e-10 * normb, window=10,
                              min_factor=2.0 if kind == "zero" else 1e6)
    x = xd.numpy()
    rel = math.sqrt(oracle.ddot1(b - oracle.spmv(A, x))) / normb
    assert res["fell_back"], res
    assert rel <= 2e-10, rel


def test_solver_graph_caches_survive_matrix_reuploads(ctx):
    """Regression: captured graphs hold matrix and workspace addresses; re-uploading matrices of
    other sizes and allocating fresh vectors (which may land on freed addresses) must never
    replay a graph built for the old buffers.  Identity network: FCG = plain CG."""
    w = np.zeros(oracle.NW1 + oracle.NW2 + oracle.NW3, np.float32)
    c = 13
    w[0 * 27 + c], w[1 * 27 + c] = 1, -1
    w[oracle.NW1 + (0 * 16 + 0) * 27 + c] = w[oracle.NW1 + (1 * 16 + 1) * 27 + c] = 1
    w[oracle.NW1 + oracle.NW2 + 0 * 27 + c], w[oracle.NW1 + oracle.NW2 + 1 * 27 + c] = 1, -1
    ctx.cnn_set_weights(w)
    for n in (12, 16, 12, 20, 16):
        A, xs, b = _system(n, 40 + n)
        normb = math.sqrt(oracle.ddot1(b))
        ctx.upload(A.rowindex, A.colindex, A.values)
        for solver in ("fp64", "mixed", "fcg"):
            bd = hpccg_amd.DeviceArray.from_numpy(ctx, b)
            xd = hpccg_amd.DeviceArray.from_numpy(ctx, np.zeros(A.n))
            if solver == "fp64":
                ctx.cg_dev(bd, xd, max_iter=500, tolerance=1e-10 * normb)
            elif solver == "mixed":
                ctx.cg_mixed_dev(bd, xd, max_iter=500, tolerance=1e-10 * normb)
            else:
                res = ctx.fcg_learned_dev((n, n, n), bd, xd, max_iter=500, tolerance=1e-10 * normb)
                assert not res["fell_back"]
            rel = math.sqrt(oracle.ddot1(b - oracle.spmv(A, xd.numpy()))) / normb
            assert rel <= 2e-10, (n, solver, rel)
            del bd, xd
