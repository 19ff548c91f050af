"""GPU tests of the pruner's kernels: the matrix-core GEMM against numpy,
and the training forward/backward against the float64 MLP oracle."""

import sys

import numpy
import pytest
import torch

from oracle import mlp_oracle as mo
from tmol_amd.generator import generate_poses
from tmol_amd.pose import RaggedPoseStack
from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _bf16_round(a):
    return torch.from_numpy(numpy.asarray(a, dtype=numpy.float32)).to(torch.bfloat16).float().numpy()


@pytest.mark.parametrize("case", ["nt_silu", "nt_f32", "kn_dsilu", "kn_f32", "km_kn_splitk"])
def test_gemm_against_numpy(case):
    from tmol_amd import prune

    _say("gemm " + case)
    rng = numpy.random.default_rng(3)
    M, N, K = 200, 96, 136
    A = _bf16_round(rng.normal(0, 1, (M, K)))
    Bnk = _bf16_round(rng.normal(0, 1, (N, K)))
    bias = rng.normal(0, 1, N).astype(numpy.float32)
    Zin = rng.normal(0, 1, (M, N)).astype(numpy.float32)
    tb = lambda a: torch.from_numpy(numpy.ascontiguousarray(a)).to(torch.bfloat16).view(torch.int16).to(DEV)  # noqa: E731
    tf = lambda a: torch.from_numpy(numpy.ascontiguousarray(a, dtype=numpy.float32)).to(DEV)  # noqa: E731
    ref = A.astype(numpy.float64) @ Bnk.T.astype(numpy.float64)
    sig = lambda z: 1 / (1 + numpy.exp(-z))  # noqa: E731
    if case == "nt_silu":
        Z = torch.empty((M, N), device=DEV)
        H = torch.empty((M, N), dtype=torch.int16, device=DEV)
        prune._gemm(1, 0, 0, tb(A), tb(Bnk), M, N, K, K, K, N, bias=tf(bias), Z=Z, H=H)
        z = ref + bias
        numpy.testing.assert_allclose(Z.cpu().numpy(), z, rtol=1e-5, atol=1e-4)
        h = H.view(torch.bfloat16).float().cpu().numpy()
        numpy.testing.assert_allclose(h, z * sig(z), rtol=1e-2, atol=1e-2)
    elif case == "nt_f32":
        C = torch.empty((M, N), device=DEV)
        prune._gemm(0, 0, 0, tb(A), tb(Bnk), M, N, K, K, K, N, C=C)
        numpy.testing.assert_allclose(C.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    elif case in ("kn_dsilu", "kn_f32"):
        Bkn = numpy.ascontiguousarray(Bnk.T)  # [K, N]
        if case == "kn_f32":
            C = torch.empty((M, N), device=DEV)
            prune._gemm(0, 0, 1, tb(A), tb(Bkn), M, N, K, K, N, N, C=C)
            numpy.testing.assert_allclose(C.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
        else:
            D = torch.empty((M, N), device=DEV)
            H = torch.empty((M, N), dtype=torch.int16, device=DEV)
            prune._gemm(2, 0, 1, tb(A), tb(Bkn), M, N, K, K, N, N, Z=tf(Zin), H=H, D=D)
            s = sig(Zin.astype(numpy.float64))
            numpy.testing.assert_allclose(D.cpu().numpy(), ref * s * (1 + Zin * (1 - s)),
                                          rtol=1e-4, atol=1e-4)
    else:
        Akm = numpy.ascontiguousarray(A.T)   # [K, M]
        Bkn = numpy.ascontiguousarray(Bnk.T)  # [K, N]
        C = torch.zeros((M, N), device=DEV)
        prune._gemm(0, 1, 1, tb(Akm), tb(Bkn), M, N, K, M, N, N, C=C, split_k=3)
        numpy.testing.assert_allclose(C.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)


def test_training_step_matches_oracle():
    """Loss, logits and every parameter gradient of the GPU training step
    against the float64 oracle on the same bf16 inputs and bf16 weights."""
    from tmol_amd import prune
    from tmol_amd.scorer import LJLKRaggedScorer

    _say("prune training step vs oracle")
    poses = generate_poses(4, (30, 80), seed=11)
    st = RaggedPoseStack(poses)
    sc = LJLKRaggedScorer(st, LJLKEnergyTerm(device=DEV), device=DEV)
    x = st.coords.to(DEV)
    Ew, labels = prune.pair_labels(sc, x)
    pairs = prune.work_list_pairs(sc)
    assert pairs.shape[0] == labels.shape[0]
    inp = prune.PruneInputs(sc)
    inp.update(x)
    params = prune.PrunerParams(DEV, seed=5)
    for k in ("W1", "W2", "W3", "emb"):  # oracle and GPU see the same bf16 values
        params.t[k].copy_(params.t[k].to(torch.bfloat16).float())
    X = prune.pair_features(inp, params, pairs)
    loss, logits, g = prune.mlp_loss_and_grads(params, X, pairs, inp.emb_row, labels, pos_weight=0.7)
    n = pairs.shape[0]
    P = {k: v.double().cpu().numpy() for k, v in params.t.items()}
    base = X.view(torch.bfloat16).double().cpu().numpy()
    rows = inp.emb_row.long().cpu().numpy()
    pr = pairs.long().cpu().numpy()
    lo, go, lgo = mo.loss_and_grads(P, base, rows[pr[:, 0]], rows[pr[:, 1]],
                                    labels.double().cpu().numpy(), pos_weight=0.7)
    _say("n=%d loss gpu %.6g oracle %.6g" % (n, float(loss), lo))
    assert abs(float(loss) - lo) <= 2e-3 * abs(lo)
    numpy.testing.assert_allclose(logits.double().cpu().numpy(), lgo, rtol=0, atol=2e-2 * max(1, abs(lgo).max()))
    for k in go:
        a = g[k].double().cpu().numpy()
        b = go[k]
        rel = numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-12)
        _say("grad %s rel err %.3g" % (k, rel))
        assert rel < 3e-2, k


def _small_setup(seed=11):
    from tmol_amd import prune
    from tmol_amd.scorer import LJLKRaggedScorer

    poses = generate_poses(4, (30, 80), seed=seed)
    st = RaggedPoseStack(poses)
    sc = LJLKRaggedScorer(st, LJLKEnergyTerm(device=DEV), device=DEV)
    x = st.coords.to(DEV)
    return prune, sc, x


def test_fused_inference_matches_training_forward():
    """The fused inference kernel's logits equal the layer-by-layer training
    forward on the same pairs (same bf16 rounding points), and compaction
    keeps exactly the pairs whose logit is at or below the cut."""
    prune, sc, x = _small_setup()
    _say("fused inference vs training forward")
    params = prune.PrunerParams(DEV, seed=9)
    pr = prune.Pruner(sc, params, threshold=0.5)
    sc.neighbors(x)
    pairs = prune.work_list_pairs(sc)
    pr.inputs.update(x)
    X = prune.pair_features(pr.inputs, params, pairs)
    _, ref_logits, _ = prune.mlp_loss_and_grads(params, X, pairs, pr.inputs.emb_row, None, train=False)
    pr.prune(x, keep_logits=True)
    n = pr.n_pairs_host()
    assert n == pairs.shape[0]
    torch.testing.assert_close(pr.logits[:n], ref_logits, rtol=0, atol=2e-2)
    # compaction: rows now hold exactly the kept pairs, in order
    gap = _sphere_gap(pr, pairs)
    keep = (pr.logits[:n] <= pr.logit_cut) | (pairs[:, 0] == pairs[:, 1]) | (gap < pr.min_gap)
    want = pairs[keep].cpu()
    got = prune.work_list_pairs(sc).cpu()
    assert torch.equal(got, want)
    assert 0 < got.shape[0] < n


def _sphere_gap(pr, pairs):
    """|c1 - c2| - r1 - r2 of each pair from the neighbor pass's spheres."""
    sph = pr.inputs.spheres()
    a, b = sph[pairs[:, 0].long()], sph[pairs[:, 1].long()]
    d = a[:, :3] - b[:, :3]
    return torch.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) - a[:, 3] - b[:, 3]


def test_gap_guard_keeps_pairs_that_could_clash():
    """With a threshold that lets the network remove every pair, the kept
    work list is exactly the self pairs and the pairs whose sphere gap (a
    lower bound on their atom-pair distances) is below the guard."""
    prune, sc, x = _small_setup(seed=13)
    _say("gap guard")
    params = prune.PrunerParams(DEV, seed=4)
    pr = prune.Pruner(sc, params, threshold=0.0)
    assert pr.min_gap == prune.MIN_GAP == 3.0
    sc.neighbors(x)
    pairs = prune.work_list_pairs(sc)
    pr.prune(x)
    gap = _sphere_gap(pr, pairs)
    assert not bool(((gap - pr.min_gap).abs() < 1e-4).any())  # nothing on the boundary
    want = pairs[(pairs[:, 0] == pairs[:, 1]) | (gap < pr.min_gap)].cpu()
    got = prune.work_list_pairs(sc).cpu()
    assert torch.equal(got, want)
    assert want.shape[0] > int((pairs[:, 0] == pairs[:, 1]).sum())


def test_pruning_nothing_reproduces_exact_scores():
    prune, sc, x = _small_setup(seed=12)
    _say("prune nothing = exact")
    params = prune.PrunerParams(DEV, seed=3)
    e0, g0 = [t.clone() for t in sc.energy_and_grad(x)]
    pr = prune.Pruner(sc, params, threshold=1.0)
    e1, g1 = pr.energy_and_grad(x)
    torch.testing.assert_close(e1, e0, rtol=1e-6, atol=1e-4)
    torch.testing.assert_close(g1, g0, rtol=1e-5, atol=1e-4)
