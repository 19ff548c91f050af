"""Learned block-pair pruner: host side.

A small MLP (40 -> 128 -> 128 -> 128 -> 1, SiLU; BF16 on the matrix cores,
FP32 accumulation) predicts, for every neighbor block pair of the work list,
P(|E_pair| < 1e-3 kcal/mol) with E_pair = 1.0 atr + 0.55 rep + 1.0 lk of the
pair.  It is trained on pair energies from the exact kernel by the library's
own forward and backward kernels; pairs predicted negligible are removed from
the work list before the scoring kernel runs; the exact kernel judges the
pruned energies and the caller falls back to exact scoring when the error
bound is exceeded.

Input features of a pair (g1, g2), columns of the 48-wide (40 used) row:
  0 centroid distance / 10 A; 1, 2 sphere radii / 10 A; 3..10 and 11..18
  the 8-dim residue-type embeddings of g1 and g2; 19..28 and 29..38 sin and
  cos of phi, psi, omega, chi1, chi2 of g1 and g2 (0, 0 where a torsion does
  not exist); 39 min(|i - j|, 32) / 32.
"""

import ctypes
import math
from typing import Dict, Optional

import numpy
import torch

from . import ops
from .pose import RaggedPoseStack

N_IN, K1, HIDDEN, EMB = 40, 48, 128, 8
TORSIONS = ("phi", "psi", "omega", "chi1", "chi2")
NEGLIGIBLE = 1e-3  # kcal/mol
WEIGHTS = (1.0, 0.55, 1.0)

_bound = False


class CPruneBlocks(ctypes.Structure):
    _fields_ = [("sphere", ctypes.c_void_p), ("sincos", ctypes.c_void_p),
                ("emb_row", ctypes.c_void_p), ("seqpos", ctypes.c_void_p),
                ("emb", ctypes.c_void_p)]


def _lib():
    global _bound
    L = ops.lib()
    if not _bound:
        VP, I32, I64, F32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
        L.tmol_ljlk_prune_block_torsions.argtypes = [I32, VP, VP, I32, VP, VP]
        L.tmol_ljlk_prune_pair_features.argtypes = [I64, VP, ctypes.POINTER(CPruneBlocks), VP, VP]
        L.tmol_ljlk_prune_gemm.argtypes = [I32, I32, I32, VP, VP, I32, I32, I32, I32, I32, I32,
                                           I32, VP, VP, VP, VP, VP, VP]
        L.tmol_ljlk_prune_head.argtypes = [I32, VP, VP, VP, VP, VP, VP, F32, I32, VP, VP, VP, VP,
                                           VP, VP, VP]
        L.tmol_ljlk_prune_colsum.argtypes = [I32, I32, I32, VP, VP, VP]
        L.tmol_ljlk_prune_emb_grad.argtypes = [I64, VP, VP, VP, VP, VP]
        L.tmol_ljlk_spheres.argtypes = [ctypes.POINTER(ops.CStack), ctypes.POINTER(ops.CBlockTypes),
                                        VP, I64, I32, ctypes.POINTER(VP)]
        _bound = True
    return L


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------------------
# topology-only inputs


def block_torsion_atoms(stack: RaggedPoseStack) -> torch.Tensor:
    """Global atom indices [G, 5, 4] of phi, psi, omega, chi1, chi2 of every
    block (-1 where the torsion does not exist: chain ends, no side chain)."""
    G = stack.n_blocks_total
    out = numpy.full((G, len(TORSIONS), 4), -1, dtype=numpy.int32)
    off = stack.block_atom_offset.numpy()
    g = 0
    for pose in stack.poses:
        rts = pose.residue_types
        n = len(rts)
        for i, rt in enumerate(rts):
            prev_ok = i > 0 and pose.connected_to_prev[i]
            next_ok = i + 1 < n and pose.connected_to_prev[i + 1]

            def resolve(ref):
                if "atom" in ref:
                    a = rt.atom_index.get(ref["atom"])
                    return -1 if a is None else int(off[g] + a)
                conn, sep = ref["connection"], ref.get("bond_sep_from_conn", 0)
                if conn == "down" and prev_ok:
                    other, og = rts[i - 1], g - 1
                    name = "C" if sep == 0 else "CA"
                elif conn == "up" and next_ok:
                    other, og = rts[i + 1], g + 1
                    name = "N" if sep == 0 else "CA"
                else:
                    return -1
                a = other.atom_index.get(name)
                return -1 if a is None else int(off[og] + a)

            for t, tname in enumerate(TORSIONS):
                tor = next((x for x in rt.torsions if x["name"] == tname), None)
                if tor is None:
                    continue
                idx = [resolve(tor[k]) for k in "abcd"]
                if min(idx) >= 0:
                    out[g, t] = idx
            g += 1
    return torch.from_numpy(out)


def embedding_rows(stack: RaggedPoseStack, vocab: Dict[str, int]) -> torch.Tensor:
    """Embedding row of every block: its residue type's base name (termini
    variants share the row of their residue)."""
    rows = []
    for pose in stack.poses:
        for rt in pose.residue_types:
            rows.append(vocab[rt.name3])
    return torch.tensor(rows, dtype=torch.int32)


def sequence_positions(stack: RaggedPoseStack) -> torch.Tensor:
    pbs = stack.pose_block_start.to(torch.int64)
    g = torch.arange(stack.n_blocks_total, dtype=torch.int64)
    return (g - pbs[stack.block_pose.to(torch.int64)]).to(torch.int32)


CANONICAL = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
             "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"]
VOCAB = {n: i for i, n in enumerate(CANONICAL)}


# ---------------------------------------------------------------------------
# parameters


# Pairs whose sphere gap |c1 - c2| - r1 - r2 (a lower bound on every
# atom-pair distance between the two blocks) is below this are always scored:
# the network never removes a pair whose atoms could come within 3 A, where a
# clash is worth orders of magnitude more than the error bound.  On the
# 1024 x 300-residue batch the guard keeps 2% of the pairs the network would
# prune.  0 disables it.
MIN_GAP = 3.0

class PrunerParams:
    """FP32 master parameters (torch tensors on the device); bf16 copies are
    made for the matrix-core kernels."""

    def __init__(self, device, seed=0, n_emb=len(CANONICAL)):
        g = torch.Generator().manual_seed(seed)

        def lin(o, i, fan_in):
            return (torch.rand((o, i), generator=g) * 2 - 1) * math.sqrt(6.0 / fan_in) * 0.5

        self.t = {
            "W1": lin(HIDDEN, K1, N_IN), "b1": torch.zeros(HIDDEN),
            "W2": lin(HIDDEN, HIDDEN, HIDDEN), "b2": torch.zeros(HIDDEN),
            "W3": lin(HIDDEN, HIDDEN, HIDDEN), "b3": torch.zeros(HIDDEN),
            "w4": lin(1, HIDDEN, HIDDEN).view(HIDDEN), "b4": torch.zeros(1),
            "emb": torch.randn((n_emb, EMB), generator=g) * 0.5,
        }
        self.t["W1"][:, N_IN:] = 0.0
        self.t = {k: v.to(device=device, dtype=torch.float32).contiguous() for k, v in self.t.items()}
        self.device = torch.device(device)

    def tensors(self):
        return list(self.t.values())

    def bf16(self, name) -> torch.Tensor:
        return self.t[name].to(torch.bfloat16).view(torch.int16)

    def state_dict(self):
        return {k: v.detach().cpu() for k, v in self.t.items()}

    def load_state_dict(self, d):
        for k, v in d.items():
            self.t[k].copy_(v.to(self.device))


# ---------------------------------------------------------------------------
# block inputs of a stack


class PruneInputs:
    """Per-block pruner inputs of a ragged stack scored by ``scorer``
    (float32): spheres from the scorer's neighbor pass, torsions, embedding
    rows, sequence positions."""

    def __init__(self, scorer, vocab=VOCAB):
        if scorer.dtype != torch.float32:
            raise ops.LJLKLibraryError("the pruner runs on float32 stacks")
        self.scorer = scorer
        dev = scorer.device
        st = scorer.stack
        self.G = st.n_blocks_total
        self.tors_atoms = block_torsion_atoms(st).to(dev).contiguous()
        self.emb_row = embedding_rows(st, vocab).to(dev)
        self.seqpos = sequence_positions(st).to(dev)
        self.sincos = torch.empty((self.G, 10), dtype=torch.float32, device=dev)

    def spheres(self) -> torch.Tensor:
        """View of the workspace spheres [G, 4] (valid after a neighbor pass)."""
        sc = self.scorer
        ptr = ctypes.c_void_p()
        ops._check(_lib().tmol_ljlk_spheres(ctypes.byref(sc.cstack), ctypes.byref(sc.bt.c),
                                            _p(sc.ws.buf), sc.ws.capacity, sc.dcode,
                                            ctypes.byref(ptr)), "spheres")
        off = ptr.value - sc.ws.buf.data_ptr()
        return sc.ws.buf[off: off + 16 * self.G].view(torch.float32).view(self.G, 4)

    def update(self, coords: torch.Tensor):
        """Torsion features at ``coords`` (after the neighbor pass)."""
        ops._check(_lib().tmol_ljlk_prune_block_torsions(
            self.G, _p(self.tors_atoms), _p(coords), ops.F32, _p(self.sincos),
            ops._stream(coords.device)), "block torsions")

    def blocks(self, params: PrunerParams) -> CPruneBlocks:
        self._sph = self.spheres()
        b = CPruneBlocks()
        b.sphere = self._sph.data_ptr()
        b.sincos = self.sincos.data_ptr()
        b.emb_row = self.emb_row.data_ptr()
        b.seqpos = self.seqpos.data_ptr()
        b.emb = params.t["emb"].data_ptr()
        return b


def pair_features(inputs: PruneInputs, params: PrunerParams, pairs: torch.Tensor) -> torch.Tensor:
    """Feature rows [n, 48] (bf16 bits as int16) of block pairs [n, 2]."""
    n = pairs.shape[0]
    X = torch.empty((n, K1), dtype=torch.int16, device=pairs.device)
    b = inputs.blocks(params)
    ops._check(_lib().tmol_ljlk_prune_pair_features(n, _p(pairs), ctypes.byref(b), _p(X),
                                                    ops._stream(pairs.device)), "pair features")
    return X


# ---------------------------------------------------------------------------
# training step (forward and backward kernels)


def _gemm(epi, a_km, b_kn, A, B, M, N, K, lda, ldb, ldc, C=None, bias=None, Z=None, H=None, D=None,
          split_k=1):
    ops._check(_lib().tmol_ljlk_prune_gemm(
        epi, a_km, b_kn, _p(A), _p(B), M, N, K, lda, ldb, ldc, split_k, _p(C), _p(bias), _p(Z),
        _p(H), _p(D), ops._stream(A.device)), "prune gemm")


def mlp_loss_and_grads(params: PrunerParams, X: torch.Tensor, pairs: torch.Tensor,
                       emb_row: torch.Tensor, labels: Optional[torch.Tensor], pos_weight=1.0,
                       train=True, sample_weights: Optional[torch.Tensor] = None):
    """Forward (and with ``train`` the backward) of the MLP on feature rows X
    [n, 48] (bf16 bits).  Returns (summed loss or None, logits [n], grads)."""
    dev = X.device
    n = X.shape[0]
    f32 = dict(dtype=torch.float32, device=dev)
    i16 = dict(dtype=torch.int16, device=dev)
    Z = [torch.empty((n, HIDDEN), **f32) for _ in range(3)]
    H = [torch.empty((n, HIDDEN), **i16) for _ in range(3)]
    W = [params.bf16(k) for k in ("W1", "W2", "W3")]
    b = [params.t[k] for k in ("b1", "b2", "b3")]
    _gemm(1, 0, 0, X, W[0], n, HIDDEN, K1, K1, K1, HIDDEN, bias=b[0], Z=Z[0], H=H[0])
    _gemm(1, 0, 0, H[0], W[1], n, HIDDEN, HIDDEN, HIDDEN, HIDDEN, HIDDEN, bias=b[1], Z=Z[1], H=H[1])
    _gemm(1, 0, 0, H[1], W[2], n, HIDDEN, HIDDEN, HIDDEN, HIDDEN, HIDDEN, bias=b[2], Z=Z[2], H=H[2])
    logits = torch.empty(n, **f32)
    if not train:
        ops._check(_lib().tmol_ljlk_prune_head(n, _p(H[2]), None, _p(params.t["w4"]),
                                               _p(params.t["b4"]), None, None, 1.0, 0, _p(logits),
                                               None, None, None, None, None, ops._stream(dev)),
                   "head")
        return None, logits, None
    g = {k: torch.zeros_like(v) for k, v in params.t.items()}
    loss = torch.zeros(1, **f32)
    D = [torch.empty((n, HIDDEN), **f32) for _ in range(3)]
    Db = [torch.empty((n, HIDDEN), **i16) for _ in range(3)]
    ops._check(_lib().tmol_ljlk_prune_head(
        n, _p(H[2]), _p(Z[2]), _p(params.t["w4"]), _p(params.t["b4"]), _p(labels),
        _p(sample_weights), float(pos_weight), 1, _p(logits), _p(loss), _p(D[2]), _p(Db[2]),
        _p(g["w4"]),
        _p(g["b4"]), ops._stream(dev)), "head")
    split = max(1, min(64, n // 2048))
    # layer 3: dW3 = dZ3^T H2, db3 = sum dZ3, dZ2 = (dZ3 W3) * silu'(Z2)
    _gemm(0, 1, 1, Db[2], H[1], HIDDEN, HIDDEN, n, HIDDEN, HIDDEN, HIDDEN, C=g["W3"], split_k=split)
    _colsum(D[2], g["b3"])
    _gemm(2, 0, 1, Db[2], W[2], n, HIDDEN, HIDDEN, HIDDEN, HIDDEN, HIDDEN, Z=Z[1], H=Db[1], D=D[1])
    _gemm(0, 1, 1, Db[1], H[0], HIDDEN, HIDDEN, n, HIDDEN, HIDDEN, HIDDEN, C=g["W2"], split_k=split)
    _colsum(D[1], g["b2"])
    _gemm(2, 0, 1, Db[1], W[1], n, HIDDEN, HIDDEN, HIDDEN, HIDDEN, HIDDEN, Z=Z[0], H=Db[0], D=D[0])
    _gemm(0, 1, 1, Db[0], X, HIDDEN, K1, n, HIDDEN, K1, K1, C=g["W1"], split_k=split)
    _colsum(D[0], g["b1"])
    dX = torch.empty((n, K1), **f32)
    _gemm(0, 0, 1, Db[0], W[0], n, K1, HIDDEN, HIDDEN, K1, K1, C=dX)
    ops._check(_lib().tmol_ljlk_prune_emb_grad(n, _p(pairs), _p(emb_row), _p(dX), _p(g["emb"]),
                                               ops._stream(dev)), "emb grad")
    g["W1"][:, N_IN:] = 0.0
    return loss, logits, g


def _colsum(D, out):
    n, c = D.shape
    ops._check(_lib().tmol_ljlk_prune_colsum(n, c, c, _p(D), _p(out), ops._stream(D.device)),
               "colsum")


def work_list_pairs(scorer) -> torch.Tensor:
    """The scorer's current work list as block pairs [n, 2] (g1, g2) on the
    device, in row order."""
    cnt, items = scorer.ws.rows(scorer.cstack, scorer.bt.c)
    cnt = cnt.to(torch.int64).clamp(max=scorer.capacity)
    mask = torch.arange(scorer.capacity, device=cnt.device)[None, :] < cnt[:, None]
    g1 = torch.arange(scorer.G, device=cnt.device, dtype=torch.int32)[:, None].expand_as(items)
    return torch.stack([g1[mask], items[mask]], dim=1).contiguous()


def pair_labels(scorer, coords, pairs_mask_order=True):
    """Weighted energy of every work-list pair (exact kernel) and the label
    |E| < 1e-3 kcal/mol, in the order of ``work_list_pairs``."""
    E = scorer.item_energies(coords)  # [3, G, cap]; also runs the neighbor pass
    cnt, _ = scorer.ws.rows(scorer.cstack, scorer.bt.c)
    cnt = cnt.to(torch.int64).clamp(max=scorer.capacity)
    mask = torch.arange(scorer.capacity, device=cnt.device)[None, :] < cnt[:, None]
    w = torch.tensor(WEIGHTS, dtype=E.dtype, device=E.device)
    Ew = (w[:, None, None] * E).sum(0)[mask]
    return Ew, (Ew.abs() < NEGLIGIBLE).to(torch.float32)


# ---------------------------------------------------------------------------
# inference: prune the work list of a scorer


def _logit(p: float) -> float:
    if p >= 1.0:
        return math.inf
    if p <= 0.0:
        return -math.inf
    return math.log(p / (1.0 - p))


class Pruner:
    """Prunes a ``LJLKRaggedScorer``'s work list in place with the MLP.

    ``threshold``: a pair is removed when the predicted probability of
    |E| < 1e-3 kcal/mol exceeds it (pairs of a block with itself are always
    kept).  All buffers are preallocated, so a neighbor pass + prune + score
    step can be captured in one hipGraph."""

    def __init__(self, scorer, params: PrunerParams, threshold: float = 0.9,
                 min_gap: float = MIN_GAP):
        L = _lib()
        VP, I32, I64, F32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
        L.tmol_ljlk_prune_scratch_bytes.argtypes = [I32]
        L.tmol_ljlk_prune_scratch_bytes.restype = I64
        L.tmol_ljlk_prune_flatten.argtypes = [I32, VP, VP, I64, VP, VP, VP, VP]
        L.tmol_ljlk_prune_infer.argtypes = [VP, VP, ctypes.POINTER(CPruneBlocks), VP, VP, VP, VP, VP,
                                            VP, VP, VP, F32, F32, VP, VP, I32, VP]
        L.tmol_ljlk_prune_compact.argtypes = [I32, VP, VP, I64, VP, VP, VP]
        self.scorer = scorer
        self.params = params
        self.inputs = PruneInputs(scorer)
        dev = scorer.device
        G, cap = scorer.G, scorer.capacity
        self.G, self.cap = G, cap
        nb = (G + 1023) // 1024
        self.scratch = torch.zeros(int(L.tmol_ljlk_prune_scratch_bytes(G)) // 8, dtype=torch.int64, device=dev)
        self.n_pairs = self.scratch[nb:nb + 1]  # device total written by flatten
        self.row_offset = torch.empty(G + 1, dtype=torch.int64, device=dev)
        self.pairs = torch.empty((G * cap, 2), dtype=torch.int32, device=dev)
        self.keep = torch.empty(G * cap, dtype=torch.uint8, device=dev)
        self.logits = torch.empty(G * cap, dtype=torch.float32, device=dev)
        self.min_gap = float(min_gap)
        self.set_threshold(threshold)
        self.refresh()

    def set_threshold(self, threshold: float):
        self.threshold = float(threshold)
        self.logit_cut = _logit(self.threshold)

    def refresh(self):
        """Copy the current parameters to the bf16 inference buffers."""
        p = self.params
        self.w = {k: p.bf16(k).contiguous() for k in ("W1", "W2", "W3")}
        self.blocks = self.inputs.blocks(p)

    def _rows(self):
        sc = self.scorer
        pc, pi = ctypes.c_void_p(), ctypes.c_void_p()
        ops._check(ops.lib().tmol_ljlk_work_list(ctypes.byref(sc.cstack), ctypes.byref(sc.bt.c),
                                                 _p(sc.ws.buf), sc.capacity, sc.dcode,
                                                 ctypes.byref(pc), ctypes.byref(pi)), "work list")
        return pc, pi

    def prune(self, coords: torch.Tensor, keep_logits: bool = False):
        """After the scorer's neighbor pass at ``coords``: features, MLP,
        compaction of the work list."""
        L = _lib()
        st = ops._stream(coords.device)
        self.inputs.update(coords)
        pc, pi = self._rows()
        ops._check(L.tmol_ljlk_prune_flatten(self.G, pc, pi, self.cap, _p(self.scratch),
                                             _p(self.row_offset), _p(self.pairs), st), "flatten")
        p = self.params.t
        ops._check(L.tmol_ljlk_prune_infer(
            _p(self.n_pairs), _p(self.pairs), ctypes.byref(self.blocks), _p(self.w["W1"]),
            _p(self.w["W2"]), _p(self.w["W3"]), _p(p["b1"]), _p(p["b2"]), _p(p["b3"]), _p(p["w4"]),
            _p(p["b4"]), self.logit_cut, self.min_gap, _p(self.keep),
            _p(self.logits) if keep_logits else None,
            0, st), "prune infer")
        ops._check(L.tmol_ljlk_prune_compact(self.G, pc, pi, self.cap, _p(self.row_offset),
                                             _p(self.keep), st), "compact")

    def compact_with(self, coords: torch.Tensor, keep: torch.Tensor):
        """After the scorer's neighbor pass: flatten and compact the work list
        with a given keep flag per flattened pair (uint8 [n]) instead of the
        network's -- the same flattening and compaction kernels as
        ``prune``."""
        L = _lib()
        st = ops._stream(coords.device)
        pc, pi = self._rows()
        ops._check(L.tmol_ljlk_prune_flatten(self.G, pc, pi, self.cap, _p(self.scratch),
                                             _p(self.row_offset), _p(self.pairs), st), "flatten")
        ops._check(L.tmol_ljlk_prune_compact(self.G, pc, pi, self.cap, _p(self.row_offset),
                                             _p(keep), st), "compact")

    def n_pairs_host(self) -> int:
        return int(self.n_pairs.item())

    def energy_and_grad(self, coords, weights=None):
        """Pruned step: neighbor pass, prune, fused energy + gradient."""
        sc = self.scorer
        w = sc._weights(weights)
        sc.neighbors(coords)
        self.prune(coords)
        sc.score(coords, ops.WGRAD, weights=w, energies=sc.energies, grad=sc.grad)
        return sc.energies, sc.grad


# ---------------------------------------------------------------------------
# training


def build_training_set(scorer, coords):
    """Pairs [n, 2], labels [n] and weighted pair energies of the scorer's
    stack at ``coords`` (exact kernel)."""
    Ew, labels = pair_labels(scorer, coords)
    pairs = work_list_pairs(scorer)
    return pairs, labels, Ew


def energy_sample_weights(Ew: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
    """Per-pair loss weights: 1 for negligible pairs, 1 + log(1 + |E| / 1e-3)
    (times ``scale``) otherwise, so that calling a high-energy pair
    negligible -- an error the whole pose total carries -- costs more."""
    a = Ew.abs().float()
    w = 1.0 + scale * torch.log1p(a / NEGLIGIBLE)
    return torch.where(a < NEGLIGIBLE, torch.ones_like(w), w).contiguous()


def train(params: PrunerParams, inputs: PruneInputs, pairs: torch.Tensor, labels: torch.Tensor,
          epochs: int = 4, batch: int = 8192, lr: float = 2e-3, pos_weight: float = 1.0, seed: int = 0,
          log=None, sample_weights: Optional[torch.Tensor] = None):
    """Adam on the fp32 master parameters with gradients from the library's
    kernels.  Returns the mean loss of each epoch."""
    opt = torch.optim.Adam(params.tensors(), lr=lr)
    n = pairs.shape[0]
    gen = torch.Generator(device="cpu").manual_seed(seed)
    history = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen).to(pairs.device)
        tot = torch.zeros((), dtype=torch.float64, device=pairs.device)
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            pb = pairs[idx].contiguous()
            yb = labels[idx].contiguous()
            wb = sample_weights[idx].contiguous() if sample_weights is not None else None
            X = pair_features(inputs, params, pb)
            loss, _, g = mlp_loss_and_grads(params, X, pb, inputs.emb_row, yb, pos_weight,
                                            sample_weights=wb)
            for k, t in params.t.items():
                t.grad = g[k]
            opt.step()
            params.t["W1"].data[:, N_IN:] = 0.0
            tot += loss.double()[0]
        history.append(float(tot) / n)
        if log:
            log("epoch %d loss %.5f" % (ep, history[-1]))
    return history


def evaluate_classifier(params, inputs, pairs, labels, batch=65536):
    """Logits of the training-path forward and accuracy at 0.5."""
    outs = []
    for b0 in range(0, pairs.shape[0], batch):
        pb = pairs[b0:b0 + batch].contiguous()
        X = pair_features(inputs, params, pb)
        _, lg, _ = mlp_loss_and_grads(params, X, pb, inputs.emb_row, None, train=False)
        outs.append(lg)
    lg = torch.cat(outs)
    acc = float(((lg > 0).float() == labels).float().mean())
    return lg, acc
