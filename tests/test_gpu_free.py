"""Free-space layer potentials on the GPU through the C ABI (the replacements of
the FMM mex functions stokesSLPfmm / stokesDLPfmm) against the oracle's direct
sums, including the stresslet identity of FMM/tests/stresslet_identity_test.m."""
import numpy as np
import pytest

import fasttools2d_amd as ft

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("Ns,Nt", [(3000, 2500), (20000, 100), (1, 5)])
def test_free_slp_matches_oracle(gpu_ctx, oracle, Ns, Nt):
    """Random points in [-1, 1]^2 (stokes_SLP_FMM_test.m:9-20); many sources and
    few targets run as source slices added in order."""
    rng = np.random.default_rng(Ns + Nt)
    src = rng.uniform(-1, 1, (Ns, 2))
    tar = rng.uniform(-1, 1, (Nt, 2))
    f = 10 * rng.random((Ns, 2))
    u1, u2 = ft.stokesSLPfmm(f[:, 0], f[:, 1], src[:, 0], src[:, 1], tar[:, 0], tar[:, 1], 0, 5,
                             ctx=gpu_ctx)
    ref = oracle.free_slp(f, src, tar)
    assert _rel(np.stack([u1, u2], 1), ref) < 1e-13


def test_free_slp_sources_as_targets(gpu_ctx, oracle):
    rng = np.random.default_rng(11)
    p = rng.uniform(-1, 1, (4000, 2))
    f = 10 * rng.random((4000, 2))
    u1, u2 = ft.stokesSLPfmm(f[:, 0], f[:, 1], p[:, 0], p[:, 1], p[:, 0], p[:, 1], 1, 3,
                             ctx=gpu_ctx)
    ref = oracle.free_slp(f, p, p, self=True)
    assert _rel(np.stack([u1, u2], 1), ref) < 1e-13
    with pytest.raises(ft.Se2pError):  # iequal = 1 with fewer targets than sources
        ft.stokesSLPfmm(f[:, 0], f[:, 1], p[:, 0], p[:, 1], p[:10, 0], p[:10, 1], 1, 3,
                        ctx=gpu_ctx)
    with pytest.raises(ft.Se2pError):  # iprec outside -2..5
        ft.stokesSLPfmm(f[:, 0], f[:, 1], p[:, 0], p[:, 1], p[:, 0], p[:, 1], 0, 9, ctx=gpu_ctx)


def test_free_dlp_matches_oracle_and_identity(gpu_ctx, oracle):
    rng = np.random.default_rng(12)
    N = 5000
    p = rng.uniform(-1, 1, (N, 2))
    f = 10 * rng.random((N, 2))
    n1 = rng.random(N)
    n = np.stack([n1, np.sqrt(1 - n1 ** 2)], 1)  # stokes_DLP_FMM_test.m:11-12
    u1, u2 = ft.stokesDLPfmm(f[:, 0], f[:, 1], p[:, 0], p[:, 1], n[:, 0], n[:, 1], ctx=gpu_ctx)
    assert _rel(np.stack([u1, u2], 1), oracle.free_dlp(f, n, p)) < 1e-13
    # stresslet identity: -u = (1, 0) inside the unit circle, 0 outside
    h = 2 * np.pi / 512
    th = h * np.arange(1, 513)
    c = np.stack([np.cos(th), np.sin(th)], 1)
    t = 2 * np.pi / 16 * np.arange(1, 17)
    ring = np.stack([np.cos(t), np.sin(t)], 1)
    pts = np.concatenate([c, 0.5 * ring, 1.5 * ring])
    ff = np.zeros((544, 2))
    ff[:512, 0] = h
    nn = np.concatenate([c, np.zeros((32, 2))])
    v1, v2 = ft.stokesDLPfmm(ff[:, 0], ff[:, 1], pts[:, 0], pts[:, 1], nn[:, 0], nn[:, 1],
                             ctx=gpu_ctx)
    assert np.abs(-v1[512:528] - 1).max() < 1e-12 and np.abs(v2[512:528]).max() < 1e-12
    assert np.abs(v1[528:]).max() < 1e-12 and np.abs(v2[528:]).max() < 1e-12
    # stokesDLP.f:31: a zero density gives zero
    z1, z2 = ft.stokesDLPfmm(0 * f[:, 0], 0 * f[:, 1], p[:, 0], p[:, 1], n[:, 0], n[:, 1],
                             ctx=gpu_ctx)
    assert not z1.any() and not z2.any()
