"""Surrogate network: the BF16 MFMA GEMM, forward/backward against the CPU
restatement, training, and what the trained network gains for the GA."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(False, True), (False, False), (True, False), (True, True)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (100, 37, 12), (257, 256, 256), (5, 300, 33)])
def test_gemm_bf16_against_float64(ta, tb, M, N, K):
    from neurogpu_amd.surrogate import from_bf16_bits, gemm_bf16, to_bf16_bits
    rng = np.random.default_rng(M * 7 + N + K)
    A = to_bf16_bits(rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32))
    B = to_bf16_bits(rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32))
    C = gemm_bf16(A, B, ta, tb, M, N, K)
    a = from_bf16_bits(A).astype(np.float64)
    b = from_bf16_bits(B).astype(np.float64)
    ref = (a.T if ta else a) @ (b.T if tb else b)
    np.testing.assert_allclose(C, ref, rtol=1e-5, atol=1e-5 * np.sqrt(K))


def test_gradients_match_cpu_restatement():
    from neurogpu_amd.surrogate import MLPEnsemble
    from oracle import mlp_ref
    rng = np.random.default_rng(1)
    ens = MLPEnsemble(12, 20, hidden=256, members=2, seed=3)
    x = rng.uniform(-1, 1, size=(300, 12)).astype(np.float32)
    t = rng.normal(size=(300, 20)).astype(np.float32)
    w = ens.get_weights()
    for mem in range(2):
        g, loss = ens.gradients(mem, x, t)
        params = ens.unpack(w, mem)
        ref, rloss = mlp_ref.gradients(params, x, t)
        assert abs(loss - rloss) < 1e-3 * abs(rloss)
        gd = ens.unpack(np.concatenate([np.zeros(mem * ens.nparam, np.float32), g]), mem)
        errs = [float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)) for a, b in zip(gd, ref)]
        print(f"member {mem}: relative gradient errors {['%.2e' % e for e in errs]}", flush=True)
        # FP32 accumulation against FP64 at the same BF16 rounding points:
        # measured at most 1e-4 relative
        for k, err in enumerate(errs):
            assert err < 1e-3, (mem, k, err)
    ens.close()


def test_training_fits_a_smooth_function():
    from neurogpu_amd.surrogate import MLPEnsemble
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(4096, 12)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.exp(x[:, 3]) - x[:, 4]], axis=1)
    ens = MLPEnsemble(12, 2, members=2, seed=0)
    losses = ens.fit(x, y, epochs=40, batch=256, lr=2e-3)
    assert losses[:, -1].max() < 0.2 * losses[:, 0].min()
    xt = rng.uniform(-1, 1, size=(1024, 12)).astype(np.float32)
    yt = np.stack([np.sin(2 * xt[:, 0]) + xt[:, 1] * xt[:, 2], np.exp(xt[:, 3]) - xt[:, 4]], axis=1)
    mu, sd = ens.predict(xt)
    r2 = 1 - ((mu - yt) ** 2).sum(0) / ((yt - yt.mean(0)) ** 2).sum(0)
    assert np.all(r2 > 0.8), r2
    assert mu.shape == (1024, 2) and np.all(sd >= 0)
    ens.close()


def test_ensemble_members_train_as_if_alone():
    """Every launch covers all members side by side (grid z = member) and
    member j's arithmetic reads only its own slots: a member trained in an
    ensemble ends with the weights it reaches trained alone from the same
    start with the same row orders (generator seed * 1009 + j), and each
    member's predictions equal those of a one-member network holding its
    weights."""
    from neurogpu_amd.surrogate import MLPEnsemble
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(1000, 12)).astype(np.float32)
    t = rng.normal(size=(1000, 3)).astype(np.float32)
    ens = MLPEnsemble(12, 3, members=3, seed=5)
    one = MLPEnsemble(12, 3, members=1, seed=6)
    P = ens.nparam
    w0 = ens.get_weights()
    one.set_weights(w0[:P])
    le = ens.train_raw(x, t, epochs=3, batch=128, lr=1e-3, seed=11)
    lo = one.train_raw(x, t, epochs=3, batch=128, lr=1e-3, seed=11)
    we, wo = ens.get_weights(), one.get_weights()
    assert np.array_equal(we[:P], wo)
    assert not np.array_equal(we[P:2 * P], w0[P:2 * P])
    np.testing.assert_allclose(le[0], lo[0], rtol=1e-5)
    xt = rng.uniform(-1, 1, size=(700, 12)).astype(np.float32)
    pe = ens.predict_raw(xt)
    for j in range(3):
        one.set_weights(we[j * P:(j + 1) * P])
        assert np.array_equal(pe[j], one.predict_raw(xt)[0])
    ens.close()
    one.close()
