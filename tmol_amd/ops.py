"""Host side of the MI355X LJ+LK op: ctypes binding of the C ABI
(include/tmol_ljlk.h), PyTorch op registration and autograd.

Public drop-in: ``ljlk_pose_scores(*27 args) -> [score, block_neighbors]``,
same arguments, shapes, dtypes and autograd behaviour as
``torch.ops.tmol_ljlk.ljlk_pose_scores`` (tmol/score/ljlk/potentials/
compiled.ops.cpp:29-279, 533-597):

* whole-pose mode: ``score`` is ``(3, P)``; the forward computes per-term
  dV/dx when ``rot_coords.requires_grad`` and the backward weights it by the
  incoming ``(3, P)`` gradient (compiled.ops.cpp:157-174);
* block-pair mode: ``score`` is ``(3, P, B, B)``; the backward re-walks the
  neighbor list with the ``(3, P, B, B)`` gradient as per-block-pair weights
  (ljlk_pose_score.impl.hh:1160-1477);
* ``block_neighbors`` is the dense ``(P, B, B)`` int32 flag matrix.

The device work is done only by the HIP library; there is no CPU path: a
CPU tensor or a missing library raises.  The primitive steps are registered
as PyTorch ops under ``torch.ops.tmol_amd``.
"""

import ctypes
import os
import threading
from typing import List, Optional

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtmol_ljlk_hip.so")

OK, EINVAL, EOVERFLOW, EHIP, ENOTSUP, ESCRATCH = 0, 1, 2, 3, 4, 5
F32, F64 = 0, 1
PADDED, RAGGED = 0, 1
DVDX, WGRAD, BLOCK_PAIR, NO_ENERGY, ACCUMULATE, GENERAL_KERNEL = 2, 4, 8, 16, 32, 128
DETERMINISTIC = 1024  # float32 DVDX / WGRAD: fixed-point global sums, bit-identical run to run
ROW_KERNEL = 4096  # energies / WGRAD: the block-pair row kernel instead of the stream kernel
WAVES_PER_SIMD, ROW_GRAB = 1 << 16, 1 << 20  # multipliers of the launch-variant fields
NBR_TILE = 64

_ERRORS = {
    EINVAL: "invalid argument",
    EOVERFLOW: "work list overflow",
    EHIP: "HIP runtime error",
    ENOTSUP: "unsupported configuration",
    ESCRATCH: "the stack needs more device scratch than the no-allocation entry point holds",
}


class LJLKLibraryError(RuntimeError):
    pass


_P32 = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p


class CStack(ctypes.Structure):
    _fields_ = [
        ("layout", ctypes.c_int32),
        ("n_poses", ctypes.c_int32),
        ("max_n_blocks", ctypes.c_int32),
        ("n_blocks_total", ctypes.c_int32),
        ("n_atoms_total", ctypes.c_int64),
        ("block_atom_offset", _VP),
        ("block_type", _VP),
        ("block_pose", _VP),
        ("pose_block_start", _VP),
        ("atom_pose", _VP),
        ("min_bondsep", _VP),
        ("inter_block_bondsep", _VP),
        ("bonded_start", _VP),
        ("bonded", _VP),
    ]


class CBlockTypes(ctypes.Structure):
    _fields_ = [
        ("n_block_types", ctypes.c_int32),
        ("max_n_atoms", ctypes.c_int32),
        ("max_n_tiles", ctypes.c_int32),
        ("max_n_conn", ctypes.c_int32),
        ("n_atom_types", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("n_atoms", _VP),
        ("n_heavy_in_tile", _VP),
        ("heavy_in_tile", _VP),
        ("atom_types", _VP),
        ("n_conn", _VP),
        ("conn_atoms", _VP),
        ("path_distance", _VP),
        ("is_ligand", _VP),
        ("type_params", _VP),
        ("global_params", _VP),
    ]


_lib = None
_lib_lock = threading.Lock()


def lib():
    """The HIP library; raises if it has not been built."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LJLKLibraryError(
                    "libtmol_ljlk_hip.so not found at %s: build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH
                )
            L = ctypes.CDLL(LIB_PATH)
            SP, BP = ctypes.POINTER(CStack), ctypes.POINTER(CBlockTypes)
            L.tmol_ljlk_count_neighbor_tiles.argtypes = [_P32, ctypes.c_int32]
            L.tmol_ljlk_count_neighbor_tiles.restype = ctypes.c_int64
            L.tmol_ljlk_fill_neighbor_tiles.argtypes = [_P32, ctypes.c_int32, _P32]
            L.tmol_ljlk_workspace_bytes.argtypes = [SP, BP, ctypes.c_int64, ctypes.c_int32]
            L.tmol_ljlk_workspace_bytes.restype = ctypes.c_int64
            L.tmol_ljlk_workspace_bytes_flags.argtypes = [SP, BP, ctypes.c_int64, ctypes.c_int32,
                                                          ctypes.c_int32]
            L.tmol_ljlk_workspace_bytes_flags.restype = ctypes.c_int64
            L.tmol_ljlk_neighbors.argtypes = [
                SP, BP, _VP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64, _VP,
                ctypes.c_int64, ctypes.c_int64, _VP, _VP,
            ]
            L.tmol_ljlk_score.argtypes = [
                SP, BP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64,
                ctypes.c_int32, _VP, _VP, _VP, _VP, ctypes.c_int32, _VP,
            ]
            L.tmol_ljlk_backward_combine.argtypes = [
                _VP, _VP, _VP, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _VP, _VP,
            ]
            L.tmol_ljlk_read_counters.argtypes = [_VP, ctypes.POINTER(ctypes.c_int64), _VP]
            L.tmol_ljlk_work_list.argtypes = [
                SP, BP, _VP, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_VP),
                ctypes.POINTER(_VP),
            ]
            L.tmol_ljlk_debug_pair_separations.argtypes = [
                SP, BP, _VP, ctypes.c_int64, ctypes.c_int32, _VP, ctypes.c_int64, _VP, _VP,
            ]
            L.tmol_ljlk_device_cu_count.argtypes = [_P32]
            L.tmol_ljlk_census.argtypes = [
                SP, BP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64, _VP, _VP,
            ]
            L.tmol_ljlk_noalloc_scratch_bytes.argtypes = []
            L.tmol_ljlk_noalloc_scratch_bytes.restype = ctypes.c_int64
            L.tmol_ljlk_score_noalloc.argtypes = [
                SP, BP, _VP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64,
                ctypes.c_int64, ctypes.c_int32, _VP, _VP, _VP, _VP, _VP, _VP,
            ]
            _lib = L
        return _lib


def _check(status, what):
    if status != OK:
        raise LJLKLibraryError("%s failed: %s (status %d)" % (what, _ERRORS.get(status, "?"), status))


def _p(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dtype_code(dt):
    if dt == torch.float32:
        return F32
    if dt == torch.float64:
        return F64
    raise LJLKLibraryError("coordinates must be float32 or float64, got %s" % dt)


def _i32(t, device):
    return t.to(device=device, dtype=torch.int32).contiguous()


def _require_device(t: torch.Tensor):
    if not t.is_cuda:
        raise LJLKLibraryError(
            "tmol_amd LJ+LK runs on ROCm devices only; got a tensor on %s" % t.device
        )


class BlockTypeTables:
    """Device copies of the block-type annotations + parameters (keeps the
    tensors alive for the C struct's pointers)."""

    def __init__(self, n_atoms, n_heavy_in_tile, heavy_in_tile, atom_types, n_conn,
                 conn_atoms, path_distance, is_ligand, type_params, global_params, dtype,
                 device):
        self.device = device
        self.t = dict(
            n_atoms=_i32(n_atoms, device),
            n_heavy_in_tile=_i32(n_heavy_in_tile, device),
            heavy_in_tile=_i32(heavy_in_tile, device),
            atom_types=_i32(atom_types, device),
            n_conn=_i32(n_conn, device),
            conn_atoms=_i32(conn_atoms, device),
            path_distance=_i32(path_distance, device),
            is_ligand=_i32(is_ligand, device),
            type_params=type_params.to(device=device, dtype=dtype).contiguous(),
            global_params=global_params.to(device=device, dtype=dtype).contiguous().view(-1),
        )
        T, Amax = self.t["atom_types"].shape
        c = CBlockTypes()
        c.n_block_types = T
        c.max_n_atoms = Amax
        c.max_n_tiles = self.t["n_heavy_in_tile"].shape[1]
        c.max_n_conn = self.t["conn_atoms"].shape[1] if self.t["conn_atoms"].dim() == 2 else 0
        c.n_atom_types = self.t["type_params"].shape[0]
        c.dtype = _dtype_code(dtype)
        for k in ("n_atoms", "n_heavy_in_tile", "heavy_in_tile", "atom_types", "n_conn",
                  "conn_atoms", "path_distance", "is_ligand", "type_params", "global_params"):
            setattr(c, k, _p(self.t[k]))
        self.c = c
        if self.t["heavy_in_tile"].shape[1] != 32 * c.max_n_tiles:
            raise LJLKLibraryError("block_type_heavy_atoms_in_tile must be [T, 32 * n_tiles]")
        if self.t["path_distance"].shape != (T, Amax, Amax):
            raise LJLKLibraryError("block_type_path_distance must be [T, Amax, Amax]")
        if self.t["type_params"].dim() != 2 or self.t["type_params"].shape[1] != 10:
            raise LJLKLibraryError("type_params must be [n_types, 10]")
        if self.t["global_params"].numel() != 5:
            raise LJLKLibraryError("global_params must hold 5 values")


_tile_cache = {}


def neighbor_tiles(pose_block_start: torch.Tensor, device) -> torch.Tensor:
    """Neighbor-pass tile list for a CSR pose layout (host computation,
    cached per layout)."""
    pbs = pose_block_start.detach().to("cpu", torch.int32).contiguous()
    key = (tuple(pbs.tolist()) if pbs.numel() < 4096 else (pbs.numel(), int(pbs.sum())), str(device))
    if key in _tile_cache and pbs.numel() < 4096:
        return _tile_cache[key]
    L = lib()
    P = pbs.numel() - 1
    pp = ctypes.cast(pbs.data_ptr(), _P32)
    n = L.tmol_ljlk_count_neighbor_tiles(pp, P)
    tiles = torch.zeros((max(n, 1), 4), dtype=torch.int32)
    _check(L.tmol_ljlk_fill_neighbor_tiles(pp, P, ctypes.cast(tiles.data_ptr(), _P32)), "fill tiles")
    tiles = tiles[:n].to(device)
    if pbs.numel() < 4096:
        _tile_cache[key] = tiles
    return tiles


class Workspace:
    """Caller-owned device workspace; ``capacity`` is the work list's row
    capacity (most neighbor blocks g2 >= g1 of any block g1)."""

    def __init__(self, cstack, cbt, capacity, dtype_code, device, flags=0, nbytes=None):
        L = lib()
        if nbytes is None:
            nbytes = L.tmol_ljlk_workspace_bytes_flags(ctypes.byref(cstack), ctypes.byref(cbt),
                                                       capacity, dtype_code, flags)
        if nbytes < 0:
            raise LJLKLibraryError("workspace size query failed")
        self.buf = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
        self.nbytes = int(nbytes)
        self.capacity = int(capacity)
        self.dtype_code = dtype_code
        self.G = cstack.n_blocks_total

    def rows(self, cstack, cbt):
        """(row_count [G] int32 view, row_items [G, capacity] int32 view)."""
        pc, pi = _VP(), _VP()
        _check(lib().tmol_ljlk_work_list(ctypes.byref(cstack), ctypes.byref(cbt), _p(self.buf),
                                         self.capacity, self.dtype_code, ctypes.byref(pc),
                                         ctypes.byref(pi)), "work list")
        base = self.buf.data_ptr()
        oc, oi = pc.value - base, pi.value - base
        G = self.G
        cnt = self.buf[oc : oc + 4 * G].view(torch.int32)
        items = self.buf[oi : oi + 4 * G * self.capacity].view(torch.int32).view(G, self.capacity)
        return cnt, items

    def pairs(self, cstack, cbt) -> torch.Tensor:
        """Work list as [n, 2] (g1, g2) pairs on the host."""
        cnt, items = self.rows(cstack, cbt)
        cnt = cnt.cpu().clamp(max=self.capacity)
        items = items.cpu()
        g1 = torch.repeat_interleave(torch.arange(self.G, dtype=torch.int32), cnt)
        mask = torch.arange(self.capacity)[None, :] < cnt[:, None]
        g2 = items[mask]
        return torch.stack([g1, g2], dim=1)

    def counters(self, device) -> List[int]:
        out = (ctypes.c_int64 * 2)()
        _check(lib().tmol_ljlk_read_counters(_p(self.buf), out, _stream(device)), "read counters")
        return [int(out[0]), int(out[1])]


def padded_row_capacity(P: int, B: int) -> int:
    # a row of a padded stack holds at most B blocks (g2 >= g1 in the pose)
    return max(B, 1)


class PaddedCall:
    """Everything one padded-layout op call needs on the device."""

    def __init__(self, args, dtype):
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
         n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
         max_n_rots_per_pose, min_bondsep, inter_block_bondsep, bt_n_atoms, bt_n_heavy_in_tile,
         bt_heavy_in_tile, bt_atom_types, bt_n_conn, bt_conn_atoms, bt_path_distance,
         bt_is_ligand, type_params, global_params, max_dis, block_pair) = args
        _require_device(coords)
        dev = coords.device
        self.device = dev
        P, B = first_rot_for_block.shape
        self.P, self.B = int(P), int(B)
        NA = coords.shape[0]
        if coords.dim() != 2 or coords.shape[1] != 3:
            raise LJLKLibraryError("rot_coords must be [n_atoms, 3]")
        if rot_coord_offset.numel() != P * B or block_type_ind_for_rot.numel() != P * B:
            raise LJLKLibraryError("rot index tensors must have P * B entries")
        if tuple(min_bondsep.shape) != (P, B, B):
            raise LJLKLibraryError("pose_stack_min_bond_separation must be [P, B, B]")
        if inter_block_bondsep.dim() != 5 or tuple(inter_block_bondsep.shape[:3]) != (P, B, B):
            raise LJLKLibraryError("pose_stack_inter_block_bondsep must be [P, B, B, C, C]")
        self.keep = dict(
            block_atom_offset=_i32(rot_coord_offset, dev),
            block_type=_i32(block_type_ind_for_rot, dev),
            atom_pose=_i32(pose_ind_for_atom, dev),
            min_bondsep=_i32(min_bondsep, dev),
            inter_block_bondsep=_i32(inter_block_bondsep, dev),
        )
        s = CStack()
        s.layout = PADDED
        s.n_poses = self.P
        s.max_n_blocks = self.B
        s.n_blocks_total = self.P * self.B
        s.n_atoms_total = NA
        for k, v in self.keep.items():
            setattr(s, k, _p(v))
        self.cstack = s
        self.bt = BlockTypeTables(bt_n_atoms, bt_n_heavy_in_tile, bt_heavy_in_tile, bt_atom_types,
                                  bt_n_conn, bt_conn_atoms, bt_path_distance, bt_is_ligand,
                                  type_params, global_params, dtype, dev)
        if self.bt.c.max_n_conn != inter_block_bondsep.shape[3]:
            raise LJLKLibraryError("inter_block_bondsep connection dims must match conn table")
        self.coords = coords.contiguous()
        self.dtype = dtype
        self.dcode = _dtype_code(dtype)
        self.max_dis = float(max_dis)
        self.block_pair = bool(block_pair)
        pbs = torch.arange(0, P + 1, dtype=torch.int32) * B
        self.tiles = neighbor_tiles(pbs, dev)
        cap = padded_row_capacity(self.P, self.B)
        self.ws = Workspace(self.cstack, self.bt.c, cap, self.dcode, dev)

    def neighbors(self, dense: Optional[torch.Tensor]):
        _check(lib().tmol_ljlk_neighbors(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), _p(self.coords), self.dcode,
            self.max_dis, _p(self.tiles), self.tiles.shape[0], _p(self.ws.buf), self.ws.nbytes,
            self.ws.capacity, _p(dense), _stream(self.device)), "ljlk neighbors")

    def score(self, flags, weights=None, energies=None, dvdx=None, grad=None, n_wave_groups=0):
        _check(lib().tmol_ljlk_score(
            ctypes.byref(self.cstack), ctypes.byref(self.bt.c), _p(self.coords), self.dcode,
            _p(self.ws.buf), self.ws.nbytes, self.ws.capacity, flags, _p(weights), _p(energies),
            _p(dvdx), _p(grad), n_wave_groups, _stream(self.device)), "ljlk score")


def _forward_impl(args, require_grad: bool):
    coords = args[0]
    dtype = coords.dtype
    call = PaddedCall(args, dtype)
    P, B = call.P, call.B
    dev = coords.device
    nbr = torch.empty((P, B, B), dtype=torch.int32, device=dev)
    call.neighbors(nbr)
    if call.block_pair:
        score = torch.empty((3, P, B, B), dtype=dtype, device=dev)
        call.score(BLOCK_PAIR, energies=score)
        return score, nbr, None, call
    score = torch.empty((3, P), dtype=dtype, device=dev)
    if require_grad:
        dvdx = torch.empty((3, coords.shape[0], 3), dtype=dtype, device=dev)
        call.score(DVDX, energies=score, dvdx=dvdx)
        return score, nbr, dvdx, call
    call.score(0, energies=score)
    return score, nbr, None, call


# ---------------------------------------------------------------------------
# PyTorch ops (primitive steps), registered under torch.ops.tmol_amd


@torch.library.custom_op("tmol_amd::ljlk_pose_forward", mutates_args=())
def ljlk_pose_forward(
    rot_coords: torch.Tensor, rot_coord_offset: torch.Tensor, pose_ind_for_atom: torch.Tensor,
    first_rot_for_block: torch.Tensor, first_rot_block_type: torch.Tensor,
    block_ind_for_rot: torch.Tensor, pose_ind_for_rot: torch.Tensor,
    block_type_ind_for_rot: torch.Tensor, n_rots_for_pose: torch.Tensor,
    rot_offset_for_pose: torch.Tensor, n_rots_for_block: torch.Tensor,
    rot_offset_for_block: torch.Tensor, max_n_rots_per_pose: int,
    pose_stack_min_bond_separation: torch.Tensor, pose_stack_inter_block_bondsep: torch.Tensor,
    block_type_n_atoms: torch.Tensor, block_type_n_heavy_atoms_in_tile: torch.Tensor,
    block_type_heavy_atoms_in_tile: torch.Tensor, block_type_atom_types: torch.Tensor,
    block_type_n_interblock_bonds: torch.Tensor,
    block_type_atoms_forming_chemical_bonds: torch.Tensor,
    block_type_path_distance: torch.Tensor, block_type_is_ligand_fragment: torch.Tensor,
    type_params: torch.Tensor, global_params: torch.Tensor, max_dis: float,
    output_block_pair_energies: bool, require_gradient: bool,
) -> List[torch.Tensor]:
    """Forward of the LJ+LK pose score: [score, block_neighbors, dV_dx]
    (dV_dx is empty unless require_gradient in whole-pose mode)."""
    args = (rot_coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
            first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
            n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
            max_n_rots_per_pose, pose_stack_min_bond_separation, pose_stack_inter_block_bondsep,
            block_type_n_atoms, block_type_n_heavy_atoms_in_tile, block_type_heavy_atoms_in_tile,
            block_type_atom_types, block_type_n_interblock_bonds,
            block_type_atoms_forming_chemical_bonds, block_type_path_distance,
            block_type_is_ligand_fragment, type_params, global_params, max_dis,
            output_block_pair_energies)
    score, nbr, dvdx, _ = _forward_impl(args, require_gradient and not output_block_pair_energies)
    if dvdx is None:
        dvdx = torch.empty((0,), dtype=rot_coords.dtype, device=rot_coords.device)
    return [score, nbr, dvdx]


@ljlk_pose_forward.register_fake
def _(rot_coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block, *rest):
    P, B = first_rot_for_block.shape
    block_pair = rest[-2]
    require = rest[-1]
    dt = rot_coords.dtype
    score = rot_coords.new_empty((3, P, B, B) if block_pair else (3, P), dtype=dt)
    nbr = rot_coords.new_empty((P, B, B), dtype=torch.int32)
    dv = rot_coords.new_empty((3, rot_coords.shape[0], 3) if (require and not block_pair) else (0,), dtype=dt)
    return [score, nbr, dv]


@torch.library.custom_op("tmol_amd::ljlk_backward_combine", mutates_args=())
def ljlk_backward_combine(dV_dx: torch.Tensor, grad_score: torch.Tensor,
                          pose_ind_for_atom: torch.Tensor) -> torch.Tensor:
    """grad[a] = sum_t grad_score[t, pose(a)] * dV_dx[t, a] (hand-written kernel)."""
    _require_device(dV_dx)
    dev = dV_dx.device
    NA = dV_dx.shape[1]
    grad = torch.empty((NA, 3), dtype=dV_dx.dtype, device=dev)
    gs = grad_score.to(dV_dx.dtype).contiguous()
    api = _i32(pose_ind_for_atom, dev)
    _check(lib().tmol_ljlk_backward_combine(_p(dV_dx.contiguous()), _p(gs), _p(api), NA,
                                            gs.shape[1], _dtype_code(dV_dx.dtype), _p(grad),
                                            _stream(dev)), "backward combine")
    return grad


@ljlk_backward_combine.register_fake
def _(dV_dx, grad_score, pose_ind_for_atom):
    return dV_dx.new_empty((dV_dx.shape[1], 3))


@torch.library.custom_op("tmol_amd::ljlk_block_pair_backward", mutates_args=())
def ljlk_block_pair_backward(
    grad_score: torch.Tensor, rot_coords: torch.Tensor, rot_coord_offset: torch.Tensor,
    pose_ind_for_atom: torch.Tensor, first_rot_for_block: torch.Tensor,
    first_rot_block_type: torch.Tensor, block_ind_for_rot: torch.Tensor,
    pose_ind_for_rot: torch.Tensor, block_type_ind_for_rot: torch.Tensor,
    n_rots_for_pose: torch.Tensor, rot_offset_for_pose: torch.Tensor,
    n_rots_for_block: torch.Tensor, rot_offset_for_block: torch.Tensor,
    max_n_rots_per_pose: int, pose_stack_min_bond_separation: torch.Tensor,
    pose_stack_inter_block_bondsep: torch.Tensor, block_type_n_atoms: torch.Tensor,
    block_type_n_heavy_atoms_in_tile: torch.Tensor,
    block_type_heavy_atoms_in_tile: torch.Tensor, block_type_atom_types: torch.Tensor,
    block_type_n_interblock_bonds: torch.Tensor,
    block_type_atoms_forming_chemical_bonds: torch.Tensor,
    block_type_path_distance: torch.Tensor, block_type_is_ligand_fragment: torch.Tensor,
    type_params: torch.Tensor, global_params: torch.Tensor, max_dis: float,
) -> torch.Tensor:
    """Block-pair backward: re-walk the neighbor list, scatter
    sum_t grad_score[t, p, b1, b2] * dV_t(b1, b2)/dx (ljlk_pose_score.impl.hh:
    1160-1477)."""
    args = (rot_coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
            first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
            n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
            max_n_rots_per_pose, pose_stack_min_bond_separation, pose_stack_inter_block_bondsep,
            block_type_n_atoms, block_type_n_heavy_atoms_in_tile, block_type_heavy_atoms_in_tile,
            block_type_atom_types, block_type_n_interblock_bonds,
            block_type_atoms_forming_chemical_bonds, block_type_path_distance,
            block_type_is_ligand_fragment, type_params, global_params, max_dis, True)
    call = PaddedCall(args, rot_coords.dtype)
    call.neighbors(None)
    grad = torch.empty_like(call.coords)
    w = grad_score.to(rot_coords.dtype).contiguous()
    call.score(BLOCK_PAIR | WGRAD | NO_ENERGY, weights=w, grad=grad)
    return grad


@ljlk_block_pair_backward.register_fake
def _(grad_score, rot_coords, *rest):
    return torch.empty_like(rot_coords)


# ---------------------------------------------------------------------------
# autograd of the primitive forward (LJLKPoseScoreOp, compiled.ops.cpp:29-279)


def _pose_forward_setup(ctx, inputs, output):
    block_pair = bool(inputs[-2])
    ctx.block_pair = block_pair
    ctx.n_inputs = len(inputs)
    if block_pair:
        # the block-pair backward re-walks the neighbor list (impl.hh:1160-1477)
        ctx.args = inputs[1:-2]
        ctx.save_for_backward(inputs[0])
    else:
        # whole pose: dV/dx [3, NA, 3] and pose_ind_for_atom (compiled.ops.cpp:109-115)
        ctx.save_for_backward(output[2], inputs[2])


def _pose_forward_backward(ctx, grads):
    grad_score = grads[0]
    none = (None,) * ctx.n_inputs
    if grad_score is None:
        return none
    if not ctx.block_pair:
        dvdx, atom_pose = ctx.saved_tensors
        if dvdx.numel() == 0:
            return none
        grad = torch.ops.tmol_amd.ljlk_backward_combine(dvdx, grad_score.contiguous(), atom_pose)
    else:
        (coords,) = ctx.saved_tensors
        grad = torch.ops.tmol_amd.ljlk_block_pair_backward(grad_score.contiguous(), coords, *ctx.args)
    return (grad,) + none[1:]


torch.library.register_autograd("tmol_amd::ljlk_pose_forward", _pose_forward_backward,
                                setup_context=_pose_forward_setup)


# ---------------------------------------------------------------------------
# The reference's op names: torch.ops.tmol_ljlk.ljlk_pose_scores and
# ljlk_rotamer_scores (TORCH_LIBRARY(tmol_ljlk, m), compiled.ops.cpp:667-670),
# with the argument names and types of ljlk_pose_scores_op /
# ljlk_rotamer_scores_op (compiled.ops.cpp:533-561, 597-625) and the Tensor[]
# return.  Like the reference's registration (a catch-all kernel around an
# autograd function), each is a composite kernel over this module's custom
# ops, whose autograd is registered with torch.library.register_autograd.
# A process holds one library per op name: this one replaces the reference's.

TMOL_LJLK_ARGS = (
    "Tensor rot_coords, Tensor rot_coord_offset, Tensor pose_ind_for_atom, "
    "Tensor first_rot_for_block, Tensor first_rot_block_type, Tensor block_ind_for_rot, "
    "Tensor pose_ind_for_rot, Tensor block_type_ind_for_rot, Tensor n_rots_for_pose, "
    "Tensor rot_offset_for_pose, Tensor n_rots_for_block, Tensor rot_offset_for_block, "
    "int max_n_rots_per_pose, Tensor pose_stack_min_bond_separation, "
    "Tensor pose_stack_inter_block_bondsep, Tensor block_type_n_atoms, "
    "Tensor block_type_n_heavy_atoms_in_tile, Tensor block_type_heavy_atoms_in_tile, "
    "Tensor block_type_atom_types, Tensor block_type_n_interblock_bonds, "
    "Tensor block_type_atoms_forming_chemical_bonds, Tensor block_type_path_distance, "
    "Tensor block_type_is_ligand_fragment, Tensor ljlk_type_params, Tensor global_params, "
    "float max_dis, bool output_block_pair_energies"
)
TMOL_LJLK_SCHEMAS = {
    "ljlk_pose_scores": "ljlk_pose_scores(%s) -> Tensor[]" % TMOL_LJLK_ARGS,
    "ljlk_rotamer_scores": "ljlk_rotamer_scores(%s) -> Tensor[]" % TMOL_LJLK_ARGS,
}


def _tmol_ljlk_pose_scores(*args):
    rot_coords, block_pair = args[0], bool(args[-1])
    require = rot_coords.requires_grad and not block_pair
    score, nbr, _ = torch.ops.tmol_amd.ljlk_pose_forward(*args, require)
    return [score, nbr]


def _tmol_ljlk_rotamer_scores(*args):
    score, idx = torch.ops.tmol_amd.ljlk_rotamer_forward(*args)
    return [score, idx]


_TMOL_LJLK_LIB = torch.library.Library("tmol_ljlk", "DEF")
for _name, _schema in TMOL_LJLK_SCHEMAS.items():
    _TMOL_LJLK_LIB.define(_schema)
_TMOL_LJLK_LIB.impl("ljlk_pose_scores", _tmol_ljlk_pose_scores, "CompositeImplicitAutograd")
_TMOL_LJLK_LIB.impl("ljlk_rotamer_scores", _tmol_ljlk_rotamer_scores, "CompositeImplicitAutograd")


def ljlk_pose_scores(*args):
    """``torch.ops.tmol_ljlk.ljlk_pose_scores`` (27 arguments); returns
    ``[score, block_neighbors]``."""
    if len(args) != 27:
        raise TypeError("ljlk_pose_scores takes 27 arguments (%d given)" % len(args))
    return torch.ops.tmol_ljlk.ljlk_pose_scores(*args)


def debug_pair_separations(args, items: torch.Tensor) -> torch.Tensor:
    """Count-pair separations computed by the kernel's code for padded-layout
    block pairs ``items`` ([n, 2] global block indices g = p * B + b)."""
    call = PaddedCall(args, args[0].dtype)
    call.neighbors(None)
    dev = call.device
    it = _i32(items, dev)
    out = torch.empty((it.shape[0], 32 * 32), dtype=torch.uint8, device=dev)
    _check(lib().tmol_ljlk_debug_pair_separations(
        ctypes.byref(call.cstack), ctypes.byref(call.bt.c), _p(call.ws.buf), call.ws.capacity,
        call.dcode, _p(it), it.shape[0], _p(out), _stream(dev)), "debug separations")
    return out.view(-1, 32, 32)


def padded_work_list(args):
    """Run the neighbor pass and return the work list as (g1, g2) pairs on
    the host, the dense flags and the overflow flag (parity tests)."""
    call = PaddedCall(args, args[0].dtype)
    nbr = torch.empty((call.P, call.B, call.B), dtype=torch.int32, device=call.device)
    call.neighbors(nbr)
    n, overflow = call.ws.counters(call.device)
    pairs = call.ws.pairs(call.cstack, call.bt.c)
    assert pairs.shape[0] == n or overflow
    return pairs, nbr, overflow


# ---------------------------------------------------------------------------
# rotamer-pair scoring (packing): tmol_ljlk::ljlk_rotamer_scores
# (compiled.ops.cpp:282-530, 597-663)

ROT_UNIT_BYTES = 32


class CRotamerSet(ctypes.Structure):
    _fields_ = [
        ("n_poses", ctypes.c_int32),
        ("max_n_blocks", ctypes.c_int32),
        ("n_rots", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("n_atoms_total", ctypes.c_int64),
        ("rot_coord_offset", _VP),
        ("block_ind_for_rot", _VP),
        ("block_type_ind_for_rot", _VP),
        ("n_rots_for_block", _VP),
        ("rot_offset_for_block", _VP),
        ("first_rot_block_type", _VP),
        ("min_bondsep", _VP),
        ("inter_block_bondsep", _VP),
    ]


_rot_bound = False


def _rot_lib():
    global _rot_bound
    L = lib()
    if not _rot_bound:
        RP, BP = ctypes.POINTER(CRotamerSet), ctypes.POINTER(CBlockTypes)
        L.tmol_ljlk_rotamer_workspace_bytes.argtypes = [RP, BP, ctypes.c_int32]
        L.tmol_ljlk_rotamer_workspace_bytes.restype = ctypes.c_int64
        L.tmol_ljlk_rotamer_neighbors.argtypes = [
            RP, BP, _VP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64, _VP]
        L.tmol_ljlk_rotamer_totals.argtypes = [
            RP, BP, ctypes.c_int32, _VP, ctypes.POINTER(ctypes.c_int64), _VP]
        L.tmol_ljlk_rotamer_dispatch.argtypes = [
            RP, BP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64, _VP, ctypes.c_int64,
            _VP, ctypes.c_int64, _VP]
        L.tmol_ljlk_rotamer_score.argtypes = [
            RP, BP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, _VP, ctypes.c_int64,
            ctypes.c_int64, ctypes.c_int32, _VP, _VP, _VP, ctypes.c_int32, _VP]
        L.tmol_ljlk_rotamer_census.argtypes = [
            RP, BP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, _VP, ctypes.c_int64,
            ctypes.c_int64, _VP, _VP]
        L.tmol_ljlk_rotamer_spheres.argtypes = [
            RP, BP, ctypes.c_int32, _VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP)]
        _rot_bound = True
    return L


class RotamerCall:
    """Everything one ``ljlk_rotamer_scores`` call needs on the device: the
    rotamer set and block-type tables as C structs, the workspace, and
    (after ``dispatch``) the dispatch list and work units."""

    def __init__(self, args, dtype):
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
         n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
         max_n_rots_per_pose, min_bondsep, inter_block_bondsep, bt_n_atoms, bt_n_heavy_in_tile,
         bt_heavy_in_tile, bt_atom_types, bt_n_conn, bt_conn_atoms, bt_path_distance,
         bt_is_ligand, type_params, global_params, max_dis, block_pair) = args
        self._init_rotamer_set(args[:15], block_pair, "ljlk_rotamer_scores")
        self.bt = BlockTypeTables(bt_n_atoms, bt_n_heavy_in_tile, bt_heavy_in_tile, bt_atom_types,
                                  bt_n_conn, bt_conn_atoms, bt_path_distance, bt_is_ligand,
                                  type_params, global_params, dtype, self.device)
        self._finish(self.bt.c, inter_block_bondsep, max_dis, dtype)

    def _init_rotamer_set(self, rot_args, block_pair, op_name):
        """Checks and device copies of the rotamer set (the op's first 15
        arguments) and its C struct."""
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
         n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
         max_n_rots_per_pose, min_bondsep, inter_block_bondsep) = rot_args
        _require_device(coords)
        if not block_pair:
            raise LJLKLibraryError(
                "%s computes rotamer-pair energies only "
                "(output_block_pair_energies must be True)" % op_name)
        dev = coords.device
        self.device = dev
        if coords.dim() != 2 or coords.shape[1] != 3:
            raise LJLKLibraryError("rot_coords must be [n_atoms, 3]")
        P, B = n_rots_for_block.shape
        R = rot_coord_offset.numel()
        if block_ind_for_rot.numel() != R or block_type_ind_for_rot.numel() != R:
            raise LJLKLibraryError("per-rotamer tensors must have n_rots entries")
        if tuple(rot_offset_for_block.shape) != (P, B) or tuple(first_rot_block_type.shape) != (P, B):
            raise LJLKLibraryError("per-block rotamer tensors must be [P, B]")
        if tuple(min_bondsep.shape) != (P, B, B):
            raise LJLKLibraryError("pose_stack_min_bond_separation must be [P, B, B]")
        if inter_block_bondsep.dim() != 5 or tuple(inter_block_bondsep.shape[:3]) != (P, B, B):
            raise LJLKLibraryError("pose_stack_inter_block_bondsep must be [P, B, B, C, C]")
        self.P, self.B, self.R = int(P), int(B), int(R)
        self.keep = dict(
            rot_coord_offset=_i32(rot_coord_offset, dev),
            block_ind_for_rot=_i32(block_ind_for_rot, dev),
            block_type_ind_for_rot=_i32(block_type_ind_for_rot, dev),
            n_rots_for_block=_i32(n_rots_for_block, dev),
            rot_offset_for_block=_i32(rot_offset_for_block, dev),
            first_rot_block_type=_i32(first_rot_block_type, dev),
            min_bondsep=_i32(min_bondsep, dev),
            inter_block_bondsep=_i32(inter_block_bondsep, dev),
        )
        c = CRotamerSet()
        c.n_poses, c.max_n_blocks, c.n_rots, c.reserved = self.P, self.B, self.R, 0
        c.n_atoms_total = coords.shape[0]
        for k, v in self.keep.items():
            setattr(c, k, _p(v))
        self.crs = c
        self.coords = coords.contiguous()

    def _finish(self, cbt, inter_block_bondsep, max_dis, dtype):
        """Workspace for the block-type struct cbt; empty dispatch."""
        if cbt.max_n_conn != inter_block_bondsep.shape[3]:
            raise LJLKLibraryError("inter_block_bondsep connection dims must match conn table")
        self.cbt = cbt
        self.dtype = dtype
        self.dcode = _dtype_code(dtype)
        self.max_dis = float(max_dis)
        nbytes = _rot_lib().tmol_ljlk_rotamer_workspace_bytes(
            ctypes.byref(self.crs), ctypes.byref(cbt), self.dcode)
        if nbytes < 0:
            raise LJLKLibraryError("rotamer workspace size query failed")
        self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        self.indices = None
        self.units = None
        self.n_pairs = self.n_units = 0

    def _s(self):
        return _stream(self.device)

    def neighbors(self):
        _check(_rot_lib().tmol_ljlk_rotamer_neighbors(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), _p(self.coords), self.dcode,
            self.max_dis, _p(self.ws), self.ws.numel(), self._s()), "rotamer neighbors")

    def totals(self):
        out = (ctypes.c_int64 * 2)()
        _check(_rot_lib().tmol_ljlk_rotamer_totals(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), self.dcode, _p(self.ws), out,
            self._s()), "rotamer totals")
        return int(out[0]), int(out[1])

    def dispatch(self):
        """Neighbor pass, totals (one host sync), dispatch list and units."""
        self.neighbors()
        self.n_pairs, self.n_units = self.totals()
        self.indices = torch.empty((3, self.n_pairs), dtype=torch.int32, device=self.device)
        self.units = torch.empty((max(self.n_units, 1) * ROT_UNIT_BYTES,), dtype=torch.uint8,
                                 device=self.device)
        _check(_rot_lib().tmol_ljlk_rotamer_dispatch(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), self.dcode, self.max_dis,
            _p(self.ws), self.ws.numel(), _p(self.indices), self.n_pairs, _p(self.units),
            self.n_units, self._s()), "rotamer dispatch")
        return self.indices

    def score(self, energies=None, dTdV=None, grad=None, flags=0, n_workgroups=0):
        _check(_rot_lib().tmol_ljlk_rotamer_score(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), _p(self.coords), self.dcode,
            _p(self.ws), self.ws.numel(), _p(self.units), self.n_units, self.n_pairs, flags,
            _p(dTdV), _p(energies), _p(grad), n_workgroups, self._s()), "rotamer score")

    def energies(self, out=None, n_workgroups=0):
        if out is None:
            out = torch.empty((3, self.n_pairs), dtype=self.dtype, device=self.device)
        self.score(energies=out, n_workgroups=n_workgroups)
        return out

    def backward(self, dTdV, out=None, n_workgroups=0):
        w = dTdV.to(self.dtype).contiguous()
        if out is None:
            out = torch.empty_like(self.coords)
        self.score(dTdV=w, grad=out, flags=WGRAD | NO_ENERGY, n_workgroups=n_workgroups)
        return out

    def census(self):
        """Branch census of the dispatch list (float32): counts of enumerated
        atom pairs per potential region (tmol_amd.flops.CENSUS_FIELDS)."""
        out = torch.zeros(16, dtype=torch.int64, device=self.device)
        _check(_rot_lib().tmol_ljlk_rotamer_census(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), _p(self.coords), self.dcode,
            _p(self.ws), self.ws.numel(), _p(self.units), self.n_units, self.n_pairs, _p(out),
            self._s()), "rotamer census")
        return out.cpu().tolist()

    def spheres(self):
        """(rot_spheres [R, 4], block_spheres [P, B, 4]) views of the workspace."""
        pr, pb = _VP(), _VP()
        _check(_rot_lib().tmol_ljlk_rotamer_spheres(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), self.dcode, _p(self.ws),
            ctypes.byref(pr), ctypes.byref(pb)), "rotamer spheres")
        base = self.ws.data_ptr()
        es = 8 if self.dtype == torch.float64 else 4
        o1, o2 = pr.value - base, pb.value - base
        rs = self.ws[o1: o1 + es * 4 * self.R].view(self.dtype).view(self.R, 4)
        bs = self.ws[o2: o2 + es * 4 * self.P * self.B].view(self.dtype).view(self.P, self.B, 4)
        return rs, bs


@torch.library.custom_op("tmol_amd::ljlk_rotamer_forward", mutates_args=())
def ljlk_rotamer_forward(
    rot_coords: torch.Tensor, rot_coord_offset: torch.Tensor, pose_ind_for_atom: torch.Tensor,
    first_rot_for_block: torch.Tensor, first_rot_block_type: torch.Tensor,
    block_ind_for_rot: torch.Tensor, pose_ind_for_rot: torch.Tensor,
    block_type_ind_for_rot: torch.Tensor, n_rots_for_pose: torch.Tensor,
    rot_offset_for_pose: torch.Tensor, n_rots_for_block: torch.Tensor,
    rot_offset_for_block: torch.Tensor, max_n_rots_per_pose: int,
    pose_stack_min_bond_separation: torch.Tensor, pose_stack_inter_block_bondsep: torch.Tensor,
    block_type_n_atoms: torch.Tensor, block_type_n_heavy_atoms_in_tile: torch.Tensor,
    block_type_heavy_atoms_in_tile: torch.Tensor, block_type_atom_types: torch.Tensor,
    block_type_n_interblock_bonds: torch.Tensor,
    block_type_atoms_forming_chemical_bonds: torch.Tensor,
    block_type_path_distance: torch.Tensor, block_type_is_ligand_fragment: torch.Tensor,
    ljlk_type_params: torch.Tensor, global_params: torch.Tensor, max_dis: float,
    output_block_pair_energies: bool,
) -> List[torch.Tensor]:
    """Forward of the rotamer-pair op (LJLKRotamerScoreOp, compiled.ops.cpp:
    282-530): [scores [3, n], dispatch indices [3, n] int32]."""
    args = (rot_coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
            first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
            n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
            max_n_rots_per_pose, pose_stack_min_bond_separation, pose_stack_inter_block_bondsep,
            block_type_n_atoms, block_type_n_heavy_atoms_in_tile, block_type_heavy_atoms_in_tile,
            block_type_atom_types, block_type_n_interblock_bonds,
            block_type_atoms_forming_chemical_bonds, block_type_path_distance,
            block_type_is_ligand_fragment, ljlk_type_params, global_params, max_dis,
            output_block_pair_energies)
    call = RotamerCall(args, rot_coords.dtype)
    idx = call.dispatch()
    return [call.energies(), idx]


@ljlk_rotamer_forward.register_fake
def _(rot_coords, *rest):
    ctx = torch.library.get_ctx()
    n = ctx.new_dynamic_size()
    return [rot_coords.new_empty((3, n)), rot_coords.new_empty((3, n), dtype=torch.int32)]


@torch.library.custom_op("tmol_amd::ljlk_rotamer_backward", mutates_args=())
def ljlk_rotamer_backward(
    grad_score: torch.Tensor, rot_coords: torch.Tensor, rot_coord_offset: torch.Tensor,
    pose_ind_for_atom: torch.Tensor, first_rot_for_block: torch.Tensor,
    first_rot_block_type: torch.Tensor, block_ind_for_rot: torch.Tensor,
    pose_ind_for_rot: torch.Tensor, block_type_ind_for_rot: torch.Tensor,
    n_rots_for_pose: torch.Tensor, rot_offset_for_pose: torch.Tensor,
    n_rots_for_block: torch.Tensor, rot_offset_for_block: torch.Tensor,
    max_n_rots_per_pose: int, pose_stack_min_bond_separation: torch.Tensor,
    pose_stack_inter_block_bondsep: torch.Tensor, block_type_n_atoms: torch.Tensor,
    block_type_n_heavy_atoms_in_tile: torch.Tensor,
    block_type_heavy_atoms_in_tile: torch.Tensor, block_type_atom_types: torch.Tensor,
    block_type_n_interblock_bonds: torch.Tensor,
    block_type_atoms_forming_chemical_bonds: torch.Tensor,
    block_type_path_distance: torch.Tensor, block_type_is_ligand_fragment: torch.Tensor,
    ljlk_type_params: torch.Tensor, global_params: torch.Tensor, max_dis: float,
    output_block_pair_energies: bool,
) -> torch.Tensor:
    """Backward of the rotamer-pair op (LJLKRotamerScoreDispatch::backward,
    impl.hh:1868-2126): every rotamer pair's atom-pair derivatives weighted by
    the incoming [3, n] gradient.  The dispatch list is rebuilt (it is a
    deterministic function of the inputs, in the reference's order)."""
    args = (rot_coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
            first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
            n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
            max_n_rots_per_pose, pose_stack_min_bond_separation, pose_stack_inter_block_bondsep,
            block_type_n_atoms, block_type_n_heavy_atoms_in_tile, block_type_heavy_atoms_in_tile,
            block_type_atom_types, block_type_n_interblock_bonds,
            block_type_atoms_forming_chemical_bonds, block_type_path_distance,
            block_type_is_ligand_fragment, ljlk_type_params, global_params, max_dis,
            output_block_pair_energies)
    call = RotamerCall(args, rot_coords.dtype)
    call.dispatch()
    if tuple(grad_score.shape) != (3, call.n_pairs):
        raise LJLKLibraryError("rotamer backward: gradient must be [3, %d]" % call.n_pairs)
    return call.backward(grad_score)


@ljlk_rotamer_backward.register_fake
def _(grad_score, rot_coords, *rest):
    return torch.empty_like(rot_coords)


def _rotamer_forward_setup(ctx, inputs, output):
    ctx.args = inputs[1:]
    ctx.save_for_backward(inputs[0])


def _rotamer_forward_backward(ctx, grads):
    grad_score = grads[0]
    none = (None,) * (len(ctx.args) + 1)
    if grad_score is None:
        return none
    (coords,) = ctx.saved_tensors
    grad = torch.ops.tmol_amd.ljlk_rotamer_backward(grad_score.contiguous(), coords, *ctx.args)
    return (grad,) + none[1:]


torch.library.register_autograd("tmol_amd::ljlk_rotamer_forward", _rotamer_forward_backward,
                                setup_context=_rotamer_forward_setup)


def ljlk_rotamer_scores(*args):
    """``torch.ops.tmol_ljlk.ljlk_rotamer_scores`` (27 arguments, the rotamer
    set in place of the pose stack's rotamer tensors); returns
    ``[scores [3, n], dispatch_indices [3, n] int32]`` with one column per
    rotamer pair (pose, rot1, rot2)."""
    if len(args) != 27:
        raise TypeError("ljlk_rotamer_scores takes 27 arguments (%d given)" % len(args))
    return torch.ops.tmol_ljlk.ljlk_rotamer_scores(*args)


# ---------------------------------------------------------------------------
# elec: a sibling two-body term on the same scheduler
# (tmol_elec::elec_pose_scores, tmol/score/elec/potentials/compiled.ops.cpp:
# 22-216, 526-586)


class CElecTables(ctypes.Structure):
    _fields_ = [
        ("partial_charge", _VP),
        ("inter_repr_path_distance", _P32),
        ("intra_repr_path_distance", _P32),
        ("global_params", ctypes.c_float * 5),
    ]


_elec_bound = False


def _elec_lib():
    global _elec_bound
    L = lib()
    with _lib_lock:
        if not _elec_bound:
            SP, BP, EP = (ctypes.POINTER(CStack), ctypes.POINTER(CBlockTypes),
                          ctypes.POINTER(CElecTables))
            L.tmol_elec_neighbors.argtypes = [
                SP, BP, EP, _VP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64, _VP,
                ctypes.c_int64, ctypes.c_int64, _VP, _VP,
            ]
            L.tmol_elec_score.argtypes = [
                SP, BP, EP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64,
                ctypes.c_int32, _VP, _VP, _VP, _VP, ctypes.c_int32, _VP,
            ]
            L.tmol_ljlk_elec_prepare.argtypes = [
                SP, BP, EP, _VP, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _VP,
            ]
            L.tmol_ljlk_elec_score.argtypes = [
                SP, BP, EP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64,
                ctypes.c_int32, _VP, _VP, _VP, ctypes.c_int32, _VP,
            ]
            L.tmol_elec_census.argtypes = [
                SP, BP, EP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64, _VP, _VP,
            ]
            RP = ctypes.POINTER(CRotamerSet)
            L.tmol_elec_rotamer_neighbors.argtypes = [
                RP, BP, EP, _VP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64, _VP,
            ]
            L.tmol_elec_rotamer_score.argtypes = [
                RP, BP, EP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, _VP, ctypes.c_int64,
                ctypes.c_int64, ctypes.c_int32, _VP, _VP, _VP, ctypes.c_int32, _VP,
            ]
            _elec_bound = True
    return L


_host_param_cache = {}


def _host_floats(t: torch.Tensor, n: int):
    """Host float32 values of a small parameter tensor; device tensors are
    read once per (storage, version) and cached, so repeated calls with the
    scoring module's fixed parameters do not synchronise."""
    if not t.is_cuda:
        return [float(v) for v in t.detach().reshape(-1).to(torch.float32).tolist()]
    key = (t.data_ptr(), t._version, t.dtype, t.numel())
    v = _host_param_cache.get(key)
    if v is None:
        v = [float(x) for x in t.detach().reshape(-1).to("cpu", torch.float32).tolist()]
        if len(_host_param_cache) > 64:
            _host_param_cache.clear()
        _host_param_cache[key] = v
    if len(v) != n:
        raise LJLKLibraryError("global_params must hold %d values" % n)
    return v


class ElecTypeTables:
    """Device copies of elec's block-type tables and the C structs that
    point at them (block types: n_atoms, n_conn, conn_atoms and elec's
    ligand flags; elec: charges, representative path distances, global
    parameters as host floats)."""

    def __init__(self, n_atoms, n_conn, conn_atoms, is_ligand, partial_charge, inter_rpd,
                 intra_rpd, global_params, dtype, device):
        T, Amax = partial_charge.shape
        C = conn_atoms.shape[1] if conn_atoms.dim() == 2 else 0
        for name, t in (("block_type_inter_repr_path_distance", inter_rpd),
                        ("block_type_intra_repr_path_distance", intra_rpd)):
            if tuple(t.shape) != (T, Amax, Amax):
                raise LJLKLibraryError("%s must be [T, Amax, Amax]" % name)
        if n_atoms.numel() != T or n_conn.numel() != T or is_ligand.numel() != T:
            raise LJLKLibraryError("block-type tables must have one entry per block type")
        self.t = dict(
            n_atoms=_i32(n_atoms, device),
            n_conn=_i32(n_conn, device),
            conn_atoms=_i32(conn_atoms, device),
            is_ligand=_i32(is_ligand, device),
            partial_charge=partial_charge.to(device=device, dtype=dtype).contiguous(),
            inter=_i32(inter_rpd, device),
            intra=_i32(intra_rpd, device),
        )
        b = CBlockTypes()
        b.n_block_types, b.max_n_atoms = int(T), int(Amax)
        b.max_n_tiles = (int(Amax) - 1) // 32 + 1
        b.max_n_conn = int(C)
        b.n_atom_types = 0
        b.dtype = _dtype_code(dtype)
        for k in ("n_atoms", "n_conn", "conn_atoms", "is_ligand"):
            setattr(b, k, _p(self.t[k]))
        self.cbt = b
        e = CElecTables()
        e.partial_charge = _p(self.t["partial_charge"])
        e.inter_repr_path_distance = ctypes.cast(self.t["inter"].data_ptr(), _P32)
        e.intra_repr_path_distance = ctypes.cast(self.t["intra"].data_ptr(), _P32)
        for i, v in enumerate(_host_floats(global_params, 5)):
            e.global_params[i] = v
        self.celec = e


class ElecPaddedCall:
    """Device state of one padded-layout elec op call."""

    def __init__(self, args, dtype):
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
         n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
         max_n_rots_per_pose, min_bondsep, inter_block_bondsep, bt_n_atoms, bt_partial_charge,
         bt_n_conn, bt_conn_atoms, bt_inter_rpd, bt_intra_rpd, bt_is_ligand, global_params,
         max_dis, block_pair) = args
        _require_device(coords)
        dev = coords.device
        self.device = dev
        P, B = first_rot_for_block.shape
        self.P, self.B = int(P), int(B)
        NA = coords.shape[0]
        if coords.dim() != 2 or coords.shape[1] != 3:
            raise LJLKLibraryError("rot_coords must be [n_atoms, 3]")
        if rot_coord_offset.numel() != P * B or block_type_ind_for_rot.numel() != P * B:
            raise LJLKLibraryError("rot index tensors must have P * B entries")
        if tuple(min_bondsep.shape) != (P, B, B):
            raise LJLKLibraryError("pose_stack_min_bond_separation must be [P, B, B]")
        if inter_block_bondsep.dim() != 5 or tuple(inter_block_bondsep.shape[:3]) != (P, B, B):
            raise LJLKLibraryError("pose_stack_inter_block_bondsep must be [P, B, B, C, C]")
        C = bt_conn_atoms.shape[1] if bt_conn_atoms.dim() == 2 else 0
        if inter_block_bondsep.shape[3] != C:
            raise LJLKLibraryError("inter_block_bondsep connection dims must match conn table")
        self.keep = dict(
            block_atom_offset=_i32(rot_coord_offset, dev),
            block_type=_i32(block_type_ind_for_rot, dev),
            atom_pose=_i32(pose_ind_for_atom, dev),
            min_bondsep=_i32(min_bondsep, dev),
            inter_block_bondsep=_i32(inter_block_bondsep, dev),
        )
        s = CStack()
        s.layout = PADDED
        s.n_poses, s.max_n_blocks = self.P, self.B
        s.n_blocks_total = self.P * self.B
        s.n_atoms_total = NA
        for k, v in self.keep.items():
            setattr(s, k, _p(v))
        self.cstack = s
        self.tables = ElecTypeTables(bt_n_atoms, bt_n_conn, bt_conn_atoms, bt_is_ligand,
                                     bt_partial_charge, bt_inter_rpd, bt_intra_rpd,
                                     global_params, dtype, dev)
        self.cbt = self.tables.cbt
        self.celec = self.tables.celec
        self.coords = coords.contiguous()
        self.dtype = dtype
        self.dcode = _dtype_code(dtype)
        self.max_dis = float(max_dis)
        self.block_pair = bool(block_pair)
        pbs = torch.arange(0, P + 1, dtype=torch.int32) * B
        self.tiles = neighbor_tiles(pbs, dev)
        self.ws = Workspace(self.cstack, self.cbt, padded_row_capacity(self.P, self.B),
                            self.dcode, dev)

    def neighbors(self, dense: Optional[torch.Tensor]):
        _check(_elec_lib().tmol_elec_neighbors(
            ctypes.byref(self.cstack), ctypes.byref(self.cbt), ctypes.byref(self.celec),
            _p(self.coords), self.dcode, self.max_dis, _p(self.tiles), self.tiles.shape[0],
            _p(self.ws.buf), self.ws.nbytes, self.ws.capacity, _p(dense),
            _stream(self.device)), "elec neighbors")

    def score(self, flags, weights=None, energies=None, dvdx=None, grad=None, n_wave_groups=0):
        _check(_elec_lib().tmol_elec_score(
            ctypes.byref(self.cstack), ctypes.byref(self.cbt), ctypes.byref(self.celec),
            _p(self.coords), self.dcode, _p(self.ws.buf), self.ws.nbytes, self.ws.capacity,
            flags, _p(weights), _p(energies), _p(dvdx), _p(grad), n_wave_groups,
            _stream(self.device)), "elec score")


class ElecPoseScoreFunction(torch.autograd.Function):
    """Forward/backward of the elec pose score (ElecPoseScoreOp,
    compiled.ops.cpp:22-216): whole-pose mode saves dV/dx [1, NA, 3] and
    scales it by the incoming (1, P) gradient; block-pair mode re-walks the
    neighbor list with the (1, P, B, B) gradient as per-block-pair weights
    (elec_pose_score.impl.hh:858-1125)."""

    @staticmethod
    def forward(ctx, *args):
        coords = args[0]
        dtype = coords.dtype
        call = ElecPaddedCall(args, dtype)
        P, B = call.P, call.B
        dev = coords.device
        nbr = torch.empty((P, B, B), dtype=torch.int32, device=dev)
        call.neighbors(nbr)
        ctx.block_pair = call.block_pair
        ctx.n_args = len(args)
        ctx.mark_non_differentiable(nbr)
        if call.block_pair:
            score = torch.empty((1, P, B, B), dtype=dtype, device=dev)
            call.score(BLOCK_PAIR, energies=score.view(-1))
            ctx.call = call
            return score, nbr
        score = torch.empty((1, P), dtype=dtype, device=dev)
        if coords.requires_grad:
            dvdx = torch.empty((1, coords.shape[0], 3), dtype=dtype, device=dev)
            call.score(DVDX, energies=score, dvdx=dvdx)
            ctx.save_for_backward(dvdx, args[2])
        else:
            call.score(0, energies=score)
            ctx.save_for_backward(None, args[2])
        return score, nbr

    @staticmethod
    def backward(ctx, grad_score, grad_nbr):
        if grad_score is None:
            return (None,) * ctx.n_args
        if not ctx.block_pair:
            dvdx, atom_pose = ctx.saved_tensors
            if dvdx is None:
                return (None,) * ctx.n_args
            # dE/dx[a] = dE/dV[pose(a)] * dV/dx[a] (compiled.ops.cpp:150-168)
            gs = grad_score.to(dvdx.dtype)[0].index_select(0, atom_pose.to(dvdx.device).long())
            grad = dvdx[0] * gs.unsqueeze(-1)
        else:
            call = ctx.call
            grad = torch.empty_like(call.coords)
            w = grad_score.to(call.dtype).contiguous()
            call.score(BLOCK_PAIR | WGRAD | NO_ENERGY, weights=w, grad=grad)
        return (grad,) + (None,) * (ctx.n_args - 1)


def elec_pose_scores(*args):
    """Drop-in for ``torch.ops.tmol_elec.elec_pose_scores`` (25 arguments);
    returns ``[score, block_neighbors]``: score (1, P) or (1, P, B, B)."""
    if len(args) != 25:
        raise TypeError("elec_pose_scores takes 25 arguments (%d given)" % len(args))
    score, nbr = ElecPoseScoreFunction.apply(*args)
    return [score, nbr]


class ElecRotamerCall(RotamerCall):
    """Device state of one ``elec_rotamer_scores`` call: the rotamer set,
    elec's block-type tables, the workspace and (after ``dispatch``) the
    dispatch list and work units of the LJ+LK rotamer dispatch at elec's
    max_dis."""

    def __init__(self, args, dtype):
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
         n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
         max_n_rots_per_pose, min_bondsep, inter_block_bondsep, bt_n_atoms, bt_partial_charge,
         bt_n_conn, bt_conn_atoms, bt_inter_rpd, bt_intra_rpd, bt_is_ligand, global_params,
         max_dis, block_pair) = args
        self._init_rotamer_set(args[:15], block_pair, "elec_rotamer_scores")
        self.tables = ElecTypeTables(bt_n_atoms, bt_n_conn, bt_conn_atoms, bt_is_ligand,
                                     bt_partial_charge, bt_inter_rpd, bt_intra_rpd,
                                     global_params, dtype, self.device)
        self.celec = self.tables.celec
        self._finish(self.tables.cbt, inter_block_bondsep, max_dis, dtype)

    def neighbors(self):
        _check(_elec_lib().tmol_elec_rotamer_neighbors(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), ctypes.byref(self.celec),
            _p(self.coords), self.dcode, self.max_dis, _p(self.ws), self.ws.numel(), self._s()),
            "elec rotamer neighbors")

    def score(self, energies=None, dTdV=None, grad=None, flags=0, n_workgroups=0):
        _check(_elec_lib().tmol_elec_rotamer_score(
            ctypes.byref(self.crs), ctypes.byref(self.cbt), ctypes.byref(self.celec),
            _p(self.coords), self.dcode, _p(self.ws), self.ws.numel(), _p(self.units),
            self.n_units, self.n_pairs, flags, _p(dTdV), _p(energies), _p(grad), n_workgroups,
            self._s()), "elec rotamer score")

    def energies(self, out=None, n_workgroups=0):
        if out is None:
            out = torch.empty((1, self.n_pairs), dtype=self.dtype, device=self.device)
        self.score(energies=out, n_workgroups=n_workgroups)
        return out

    def census(self):
        raise NotImplementedError("no census for elec rotamer pairs")


class ElecRotamerScoreFunction(torch.autograd.Function):
    """Forward/backward of the elec rotamer-pair op (ElecRotamerScoreOp,
    tmol/score/elec/potentials/compiled.ops.cpp:275-523): the backward weights
    every rotamer pair's atom-pair derivatives by the incoming [1, n]
    gradient (ElecRotamerScoreDispatch::backward,
    elec_pose_score.impl.hh:1437-1700)."""

    @staticmethod
    def forward(ctx, *args):
        call = ElecRotamerCall(args, args[0].dtype)
        idx = call.dispatch()
        score = call.energies()
        ctx.call = call
        ctx.n_args = len(args)
        ctx.mark_non_differentiable(idx)
        return score, idx

    @staticmethod
    def backward(ctx, grad_score, grad_idx):
        if grad_score is None:
            return (None,) * ctx.n_args
        grad = ctx.call.backward(grad_score)
        return (grad,) + (None,) * (ctx.n_args - 1)


def elec_rotamer_scores(*args):
    """Drop-in for ``torch.ops.tmol_elec.elec_rotamer_scores`` (25
    arguments, the rotamer set in place of the pose stack's rotamer
    tensors); returns ``[scores [1, n], dispatch_indices [3, n] int32]`` with
    one column per rotamer pair (pose, rot1, rot2)."""
    if len(args) != 25:
        raise TypeError("elec_rotamer_scores takes 25 arguments (%d given)" % len(args))
    score, idx = ElecRotamerScoreFunction.apply(*args)
    return [score, idx]
