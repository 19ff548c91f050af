"""Persistent ragged-batch scorer: one neighbor pass + one persistent scoring
launch over a global block-pair work list for a whole ragged stack of poses,
with preallocated buffers, optional hipGraph capture of the
energy + gradient step and an autotuned launch shape.

This is the batched path (poses of uneven length scored in one launch); the
padded drop-in op lives in ``ops.py``.
"""

import ctypes
import json
import os
import time
from typing import Optional

import torch

from . import ops
from .flops import census_dict
from .pose import RaggedPoseStack



def _tune_cache_path() -> str:
    """The autotuner's results file, in the user's cache directory
    ($TMOL_AMD_CACHE, else $XDG_CACHE_HOME/tmol_amd, else ~/.cache/tmol_amd);
    never inside the package."""
    d = os.environ.get("TMOL_AMD_CACHE")
    if not d:
        base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
        d = os.path.join(base, "tmol_amd")
    return os.path.join(d, "autotune.json")


def ragged_cstack(stack: RaggedPoseStack, dev):
    """The C stack struct of a ragged pose stack and the device tensors its
    pointers refer to."""
    i32 = lambda t: ops._i32(t, dev)  # noqa: E731
    keep = dict(
        block_atom_offset=i32(stack.block_atom_offset),
        block_type=i32(stack.block_type),
        block_pose=i32(stack.block_pose),
        pose_block_start=i32(stack.pose_block_start),
        atom_pose=i32(stack.atom_pose),
        bonded_start=i32(stack.bonded_start),
        bonded=i32(stack.bonded),
    )
    s = ops.CStack()
    s.layout = ops.RAGGED
    s.n_poses = stack.n_poses
    s.max_n_blocks = stack.max_n_blocks
    s.n_blocks_total = stack.n_blocks_total
    s.n_atoms_total = stack.n_atoms_total
    for k, v in keep.items():
        setattr(s, k, ops._p(v))
    return s, keep


class LJLKRaggedScorer:
    """Scores a ``RaggedPoseStack`` with the HIP library.

    ``energy_and_grad(coords, weights)`` is the fused step: per-pose,
    per-term energies [3, P] and the weighted gradient
    sum_t weights[t, p] dE_t/dx [NA, 3] in one scoring pass."""

    def __init__(self, stack: RaggedPoseStack, term, device="cuda", dtype=torch.float32,
                 capacity: Optional[int] = None, wave_groups: Optional[int] = None,
                 deterministic: bool = False):
        dev = torch.device(device)
        self.device = dev
        self.dtype = dtype
        self.dcode = ops._dtype_code(dtype)
        self.stack = stack
        # deterministic: gradients (and per-pose energies) summed in 64-bit
        # fixed point, so float32 results are bit-identical run to run
        # (float32 only; float64 sums are already far below its tolerance)
        self.det_flags = ops.DETERMINISTIC if (deterministic and dtype == torch.float32) else 0
        pbt = stack.packed_block_types
        self.P = stack.n_poses
        self.NA = stack.n_atoms_total
        self.G = stack.n_blocks_total
        self.cstack, self.keep = ragged_cstack(stack, dev)
        self.bt = ops.BlockTypeTables(
            pbt.n_atoms, pbt.ljlk_n_heavy_atoms_in_tile, pbt.ljlk_heavy_atoms_in_tile,
            pbt.atom_types, pbt.n_conn, pbt.conn_atom, pbt.ljlk_bond_separation,
            pbt.ljlk_is_ligand_fragment, term.type_params, term.global_params, dtype, dev,
        )
        self.max_dis = term.max_dis
        self.tiles = ops.neighbor_tiles(stack.pose_block_start, dev)
        max_nb = int(stack.n_blocks_per_pose.max())
        if capacity is None:
            # a row holds the neighbor blocks g2 >= g1 of one pose: never more
            # than the longest pose's block count, so the default cannot
            # overflow whatever the coordinates (e.g. line-search trial steps
            # that scatter atoms); a smaller capacity is checked after the
            # first neighbor pass
            capacity = max_nb
        self.capacity = int(capacity)
        # a deterministic workspace has room for the fixed-point DVDX planes
        self.ws = ops.Workspace(self.cstack, self.bt.c, self.capacity, self.dcode, dev,
                                flags=self.det_flags | (ops.DVDX if self.det_flags else 0))
        self.energies = torch.zeros((3, self.P), dtype=dtype, device=dev)
        self.grad = torch.zeros((self.NA, 3), dtype=dtype, device=dev)
        self.set_launch(wave_groups=wave_groups or 0)
        self.graph = None
        self._graph_coords = None
        self._graph_weights = None
        self._checked_capacity = capacity >= max_nb

    # -- raw steps -----------------------------------------------------------
    def neighbors(self, coords: torch.Tensor):
        ops._check(ops.lib().tmol_ljlk_neighbors(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), ops._p(coords), self.dcode,
            self.max_dis, ops._p(self.tiles), self.tiles.shape[0], ops._p(self.ws.buf),
            self.ws.nbytes, self.ws.capacity, None, ops._stream(self.device)), "neighbors")

    def score(self, coords, flags, weights=None, energies=None, dvdx=None, grad=None):
        flags = flags | self.launch_flags()
        if self.det_flags and flags & (ops.DVDX | ops.WGRAD) and not flags & ops.BLOCK_PAIR:
            flags |= self.det_flags
        ops._check(ops.lib().tmol_ljlk_score(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), ops._p(coords), self.dcode,
            ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, flags, ops._p(weights),
            ops._p(energies), ops._p(dvdx), ops._p(grad), int(self.wave_groups),
            ops._stream(self.device)), "score")

    def counters(self):
        return self.ws.counters(self.device)

    def check_capacity(self):
        n, overflow = self.counters()
        if overflow:
            raise ops.LJLKLibraryError(
                "work list overflow: %d block pairs > capacity %d" % (n, self.capacity))
        return n

    # -- fused step ------------------------------------------------------------
    def _weights(self, weights):
        if weights is None:
            weights = torch.tensor([1.0, 0.55, 1.0])
        w = torch.as_tensor(weights, dtype=self.dtype, device=self.device)
        if w.dim() == 1:
            w = w[:, None].expand(3, self.P)
        return w.contiguous()

    def energy_and_grad(self, coords: torch.Tensor, weights=None):
        """Neighbor pass + fused scoring; returns views of the scorer's
        output buffers (energies [3, P], grad [NA, 3])."""
        w = self._weights(weights)
        self.neighbors(coords)
        self.score(coords, ops.WGRAD, weights=w, energies=self.energies, grad=self.grad)
        if not self._checked_capacity:
            self.check_capacity()
            self._checked_capacity = True
        return self.energies, self.grad

    def energies_only(self, coords: torch.Tensor):
        self.neighbors(coords)
        self.score(coords, 0, energies=self.energies)
        return self.energies

    def item_energies(self, coords: torch.Tensor) -> torch.Tensor:
        """Energies of every work-list block pair, [3, G, capacity]: entry
        (t, g1, k) is term t of the pair (g1, row_items[g1, k]); entries past
        row_count[g1] are zero."""
        out = torch.empty((3, self.G, self.capacity), dtype=self.dtype, device=self.device)
        self.neighbors(coords)
        self.score(coords, ops.BLOCK_PAIR, energies=out)
        return out

    def debug_pair_separations(self, coords: torch.Tensor, items: torch.Tensor) -> torch.Tensor:
        """Count-pair separations computed by the scoring kernel's own code for
        block pairs ``items`` [n, 2] (global block indices g1 <= g2 of one
        pose): [n, 32, 32] uint8, 255 past the blocks' atoms."""
        self.neighbors(coords)
        it = ops._i32(items, self.device)
        out = torch.empty((it.shape[0], 32 * 32), dtype=torch.uint8, device=self.device)
        ops._check(ops.lib().tmol_ljlk_debug_pair_separations(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), ops._p(self.ws.buf),
            self.ws.capacity, self.dcode, ops._p(it), it.shape[0], ops._p(out),
            ops._stream(self.device)), "debug separations")
        return out.view(-1, 32, 32)

    def per_term_gradients(self, coords: torch.Tensor):
        dvdx = torch.empty((3, self.NA, 3), dtype=self.dtype, device=self.device)
        self.neighbors(coords)
        self.score(coords, ops.DVDX, energies=self.energies, dvdx=dvdx)
        return self.energies, dvdx

    # -- hipGraph ------------------------------------------------------------
    def capture(self, coords: torch.Tensor, weights=None):
        """Capture neighbors + fused scoring into a graph over static input
        buffers; ``replay(coords)`` copies new coordinates in and replays."""
        self._graph_coords = coords.detach().clone().contiguous()
        self._graph_weights = self._weights(weights).clone()
        # warm up outside capture (first launches load code objects)
        self.energy_and_grad(self._graph_coords, self._graph_weights)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.neighbors(self._graph_coords)
                self.score(self._graph_coords, ops.WGRAD, weights=self._graph_weights,
                           energies=self.energies, grad=self.grad)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self.graph = g
        return g

    def replay(self, coords: Optional[torch.Tensor] = None):
        if self.graph is None:
            raise RuntimeError("capture() first")
        if coords is not None and coords.data_ptr() != self._graph_coords.data_ptr():
            self._graph_coords.copy_(coords)
        self.graph.replay()
        return self.energies, self.grad

    # -- census ----------------------------------------------------------------
    def census(self, coords: torch.Tensor) -> dict:
        """Branch census of the current work list (float32 scorer)."""
        out = torch.zeros(16, dtype=torch.int64, device=self.device)
        self.neighbors(coords)
        ops._check(ops.lib().tmol_ljlk_census(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), ops._p(coords), self.dcode,
            ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, ops._p(out),
            ops._stream(self.device)), "census")
        return census_dict(out.cpu().tolist())

    # -- autotuner -------------------------------------------------------------
    def launch_flags(self) -> int:
        """Score-flag bits of the chosen launch variant (the kernel, waves per
        SIMD it is compiled for, rows per queue grab)."""
        return (ops.WAVES_PER_SIMD * int(self.waves_per_simd)
                | ops.ROW_GRAB * int(self.row_grab)
                | (ops.ROW_KERNEL if self.kernel == "row" else 0))

    def kernel_name(self) -> str:
        """The scoring kernel the energy + gradient step runs."""
        return "score_rows_kernel" if self.kernel == "row" else "score_stream_kernel"

    def _time_step(self, coords, w, reps):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        stream = torch.cuda.current_stream(self.device)
        self.energy_and_grad(coords, w)
        torch.cuda.synchronize(self.device)
        for a, b in ev:
            a.record(stream)
            self.energy_and_grad(coords, w)
            b.record(stream)
        torch.cuda.synchronize(self.device)
        t = sorted(a.elapsed_time(b) for a, b in ev)
        return 1e-3 * t[len(t) // 2]

    def autotune(self, coords, weights=None, reps=5, cache=True, log=None):
        """Choose the launch of the energy + gradient step by timing it on
        this batch (HIP events, median of ``reps``): the scoring kernel's
        waves per SIMD (compiled variants 3, 4, 5), the persistent grid
        (workgroups of 4 waves; 0 = one per resident slot) and the rows taken
        per queue grab, searched one after the other from the library's
        defaults.  Results are cached per device and workload signature in
        the user's cache directory.  The kernel itself is the first choice:
        the stream kernel (each row's atoms streamed through test chunks) or
        the row kernel (block pair by block pair)."""
        key = "%s|G%d|NA%d|P%d|%s|v3" % (torch.cuda.get_device_name(self.device), self.G,
                                         self.NA, self.P, self.dtype)
        table = {}
        path = _tune_cache_path()
        if cache and os.path.exists(path):
            try:
                table = json.load(open(path))
            except (OSError, ValueError):
                table = {}
        if key in table:
            self.set_launch(**table[key]["choice"])
            return table[key]
        cu = ctypes.c_int32(0)
        ops._check(ops.lib().tmol_ljlk_device_cu_count(ctypes.byref(cu)), "cu count")
        n_cu = max(int(cu.value), 1)
        w = self._weights(weights)
        choice = {"kernel": "stream", "waves_per_simd": 0, "wave_groups": 0, "row_grab": 0}
        timings = {}

        def run(**kw):
            c = dict(choice, **kw)
            self.set_launch(**c)
            t = self._time_step(coords, w, reps)
            timings["%s|simd%d|wg%d|grab%d" % (c["kernel"], c["waves_per_simd"],
                                               c["wave_groups"], c["row_grab"])] = t
            if log:
                log("autotune %s: %.4f ms" % (c, 1e3 * t))
            return t

        stages = [("kernel", ["stream", "row"]),
                  ("waves_per_simd", [0, 3, 4, 5] if self.dtype == torch.float32 else [0]),
                  ("wave_groups", [0] + [n_cu * k for k in (2, 4, 8)]),
                  ("row_grab", [0, 2, 4])]
        for name, values in stages:
            best = min(values, key=lambda v: run(**{name: v}))
            choice[name] = best if name == "kernel" else int(best)
        self.set_launch(**choice)
        entry = {"choice": choice, "cu": n_cu, "timings_s": timings,
                 "best_s": min(timings.values())}
        if cache:
            table[key] = entry
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                with open(path, "w") as f:
                    json.dump(table, f, indent=1)
            except OSError:
                pass
        return entry

    def set_launch(self, kernel="stream", waves_per_simd=0, wave_groups=0, row_grab=0):
        self.kernel = str(kernel)
        self.waves_per_simd = int(waves_per_simd)
        self.wave_groups = int(wave_groups)
        self.row_grab = int(row_grab)


class LJLKElecRaggedScorer(LJLKRaggedScorer):
    """LJ+LK and elec of a ``RaggedPoseStack`` in one pass: one neighbor pass
    (LJ+LK's 6 A reach covers elec's 5.5 A) and one launch of the scoring
    kernel with elec as a fourth term sharing staging, cull and pair tests.
    ``energy_and_grad(coords, weights)`` returns energies [4, P] (atr, rep,
    lk, elec) and sum_t weights[t, p] dE_t/dx; default weights are beta2016's
    1.0, 0.55, 1.0, 1.0."""

    def __init__(self, stack: RaggedPoseStack, ljlk_term, elec_term, device="cuda",
                 dtype=torch.float32, capacity: Optional[int] = None,
                 wave_groups: Optional[int] = None):
        super().__init__(stack, ljlk_term, device=device, dtype=dtype, capacity=capacity,
                         wave_groups=wave_groups)
        pbt = stack.packed_block_types
        t = elec_term.packed_tables(pbt)
        self.elec = ops.ElecTypeTables(
            pbt.n_atoms, pbt.n_conn, pbt.conn_atom, t["is_ligand_fragment"], t["partial_charge"],
            t["inter_repr_path_distance"], t["intra_repr_path_distance"],
            elec_term.global_params, dtype, self.device)
        self.energies = torch.zeros((4, self.P), dtype=dtype, device=self.device)
        ops._check(ops._elec_lib().tmol_ljlk_elec_prepare(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), ctypes.byref(self.elec.celec),
            ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, self.dcode,
            ops._stream(self.device)), "elec prepare")

    def score(self, coords, flags, weights=None, energies=None, dvdx=None, grad=None):
        if dvdx is not None or flags & ~(ops.WGRAD | ops.NO_ENERGY | ops.ACCUMULATE):
            raise ops.LJLKLibraryError("the fused LJ+LK+elec pass supports energies and WGRAD only")
        ops._check(ops._elec_lib().tmol_ljlk_elec_score(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), ctypes.byref(self.elec.celec),
            ops._p(coords), self.dcode, ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity,
            flags, ops._p(weights), ops._p(energies), ops._p(grad), int(self.wave_groups),
            ops._stream(self.device)), "ljlk+elec score")

    def _weights(self, weights):
        if weights is None:
            weights = torch.tensor([1.0, 0.55, 1.0, 1.0])
        w = torch.as_tensor(weights, dtype=self.dtype, device=self.device)
        if w.dim() == 1:
            w = w[:, None].expand(4, self.P)
        return w.contiguous()


class ElecRaggedScorer:
    """Elec over a ``RaggedPoseStack`` on the same scheduler: its own
    neighbor pass (elec's max_dis) into its own work list, then the
    persistent elec row kernel.  ``energy_and_grad(coords, weights)`` gives
    per-pose energies [P] and the weighted gradient
    sum_p weights[p] dE_p/dx [NA, 3] in one pass."""

    def __init__(self, stack: RaggedPoseStack, term, device="cuda", dtype=torch.float32,
                 capacity: Optional[int] = None, wave_groups: Optional[int] = None,
                 deterministic: bool = False):
        dev = torch.device(device)
        self.device = dev
        self.dtype = dtype
        self.dcode = ops._dtype_code(dtype)
        self.stack = stack
        # deterministic: gradients (and per-pose energies) summed in 64-bit
        # fixed point, so float32 results are bit-identical run to run
        # (float32 only; float64 sums are already far below its tolerance)
        self.det_flags = ops.DETERMINISTIC if (deterministic and dtype == torch.float32) else 0
        pbt = stack.packed_block_types
        self.P = stack.n_poses
        self.NA = stack.n_atoms_total
        self.G = stack.n_blocks_total
        self.cstack, self.keep = ragged_cstack(stack, dev)
        t = term.packed_tables(pbt)
        self.tables = ops.ElecTypeTables(
            pbt.n_atoms, pbt.n_conn, pbt.conn_atom, t["is_ligand_fragment"], t["partial_charge"],
            t["inter_repr_path_distance"], t["intra_repr_path_distance"], term.global_params,
            dtype, dev)
        self.max_dis = term.max_dis
        self.tiles = ops.neighbor_tiles(stack.pose_block_start, dev)
        max_nb = int(stack.n_blocks_per_pose.max())
        self.capacity = int(capacity or max_nb)
        self.ws = ops.Workspace(self.cstack, self.tables.cbt, self.capacity, self.dcode, dev)
        self.energies = torch.zeros((self.P,), dtype=dtype, device=dev)
        self.grad = torch.zeros((self.NA, 3), dtype=dtype, device=dev)
        self.wave_groups = wave_groups or 0
        self._checked_capacity = self.capacity >= max_nb

    def _refs(self):
        return ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ctypes.byref(self.tables.celec)

    def neighbors(self, coords: torch.Tensor):
        ops._check(ops._elec_lib().tmol_elec_neighbors(
            *self._refs(), ops._p(coords), self.dcode, self.max_dis, ops._p(self.tiles),
            self.tiles.shape[0], ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, None,
            ops._stream(self.device)), "elec neighbors")

    def score(self, coords, flags, weights=None, energies=None, dvdx=None, grad=None):
        ops._check(ops._elec_lib().tmol_elec_score(
            *self._refs(), ops._p(coords), self.dcode, ops._p(self.ws.buf), self.ws.nbytes,
            self.ws.capacity, flags, ops._p(weights), ops._p(energies), ops._p(dvdx),
            ops._p(grad), int(self.wave_groups), ops._stream(self.device)), "elec score")

    def check_capacity(self):
        n, overflow = self.ws.counters(self.device)
        if overflow:
            raise ops.LJLKLibraryError(
                "work list overflow: %d block pairs > capacity %d" % (n, self.capacity))
        return n

    def _weights(self, weights):
        if weights is None:
            weights = 1.0
        w = torch.as_tensor(weights, dtype=self.dtype, device=self.device)
        if w.dim() == 0:
            w = w.expand(self.P)
        return w.contiguous()

    def energy_and_grad(self, coords: torch.Tensor, weights=None):
        w = self._weights(weights)
        self.neighbors(coords)
        self.score(coords, ops.WGRAD, weights=w, energies=self.energies, grad=self.grad)
        if not self._checked_capacity:
            self.check_capacity()
            self._checked_capacity = True
        return self.energies, self.grad

    def energies_only(self, coords: torch.Tensor):
        self.neighbors(coords)
        self.score(coords, 0, energies=self.energies)
        return self.energies

    def census(self, coords: torch.Tensor) -> dict:
        """Atom-pair census of the current work list (float32 scorer)."""
        from .flops import ELEC_CENSUS_FIELDS

        out = torch.zeros(16, dtype=torch.int64, device=self.device)
        self.neighbors(coords)
        ops._check(ops._elec_lib().tmol_elec_census(
            *self._refs(), ops._p(coords), self.dcode, ops._p(self.ws.buf), self.ws.nbytes,
            self.ws.capacity, ops._p(out), ops._stream(self.device)), "elec census")
        return {k: int(v) for k, v in zip(ELEC_CENSUS_FIELDS, out.cpu().tolist())}


class RaggedScoreFunction(torch.autograd.Function):
    """Per-pose, per-term energies [3, P] of a ragged stack with a hand-written
    backward: the forward saves the per-term dV/dx [3, NA, 3] (pre-reduced
    per atom in LDS by the scoring kernel) and the backward combines it with
    the incoming [3, P] gradient in one kernel, as the reference's
    LJLKPoseScoreOp does for the padded op (compiled.ops.cpp:157-174)."""

    @staticmethod
    def forward(ctx, coords: torch.Tensor, scorer: "LJLKRaggedScorer"):
        if coords.requires_grad:
            energies, dvdx = scorer.per_term_gradients(coords.detach())
            ctx.save_for_backward(dvdx, scorer.keep["atom_pose"])
        else:
            energies = scorer.energies_only(coords)
        return energies.clone()

    @staticmethod
    def backward(ctx, grad_energies):
        dvdx, atom_pose = ctx.saved_tensors
        grad = torch.ops.tmol_amd.ljlk_backward_combine(dvdx, grad_energies.contiguous(), atom_pose)
        return grad, None


def ragged_pose_scores(scorer: "LJLKRaggedScorer", coords: torch.Tensor) -> torch.Tensor:
    """[3, P] LJ+LK energies (atr, rep, lk) of the scorer's stack at
    ``coords``, differentiable with respect to ``coords``."""
    return RaggedScoreFunction.apply(coords, scorer)

