"""lk_ball on the HIP library: device tables, the drop-in pose ops and a
ragged-batch scorer.

Drop-ins (same argument lists and outputs as the reference's ops, on device
tensors):

* ``gen_pose_waters(*33 args)`` -> waters [NA, 4, 3] (NaN where none), for
  ``tmol_lk_ball::gen_pose_waters`` (tmol/score/lk_ball/potentials/
  compiled.ops.cpp:283-353, PoseWaterGen);
* ``lk_ball_pose_score(*28 args)`` -> [score, block_neighbors], score
  [4, P] or [4, P, B, B], for ``tmol_lk_ball::lk_ball_pose_score``
  (compiled.ops.cpp:356-907, LKBallPoseScoreOp).

Both are differentiable like the reference's: the score's backward yields
gradients for the coordinates and for the waters, and the waters' backward
carries the latter to the coordinates.  The reference's block-type tables
(tile-ordered donors, acceptors, polar and occluder lists, the all-bonds
table) are converted once per table set into this library's form: per block
type an occluder list (polar atoms first) with parameters, and one water
recipe per water slot (the atoms its frame is built from).

``LKBallRaggedScorer`` scores a RaggedPoseStack: neighbor pass, waters,
energies and the weighted gradient (score, then the waters' backward).
Nothing here falls back to a CPU path: without the HIP library every entry
point raises.
"""

import ctypes
import hashlib
import threading
from typing import Optional

import numpy
import torch

from . import ops
from .score.lk_ball_energy_term import RECIPE_WIDTH, water_recipes

_VP = ctypes.c_void_p
_P32 = ctypes.POINTER(ctypes.c_int32)
MAX_OCC = 32
MAX_POLAR = 16


class CLKBallTables(ctypes.Structure):
    _fields_ = [
        ("max_n_occ", ctypes.c_int32),
        ("max_n_recipes", ctypes.c_int32),
        ("max_n_polar", ctypes.c_int32),
        ("n_occ", _P32),
        ("n_polar", _P32),
        ("occ_atoms", _P32),
        ("occ_params", _VP),
        ("n_recipes", _P32),
        ("recipes", _P32),
        ("path_distance", _P32),
        ("global_params", _VP),
        ("watergen_params", _VP),
        ("water_tors", _VP),
    ]


_bound = False
_lock = threading.Lock()


def _lkb_lib():
    global _bound
    L = ops.lib()
    with _lock:
        if not _bound:
            SP, BP, KP = (ctypes.POINTER(ops.CStack), ctypes.POINTER(ops.CBlockTypes),
                          ctypes.POINTER(CLKBallTables))
            L.tmol_lkball_waters.argtypes = [SP, BP, KP, _VP, ctypes.c_int32, _VP, _VP]
            L.tmol_lkball_waters_backward.argtypes = [SP, BP, KP, _VP, ctypes.c_int32, _VP, _VP,
                                                      _VP]
            L.tmol_lkball_neighbors.argtypes = [
                SP, BP, _VP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64, _VP,
                ctypes.c_int64, ctypes.c_int64, _VP, _VP]
            L.tmol_lkball_score.argtypes = [
                SP, BP, KP, _VP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64,
                ctypes.c_int32, _VP, _VP, _VP, _VP, ctypes.c_int32, _VP]
            L.tmol_lkball_workspace_bytes.argtypes = [SP, BP, KP, ctypes.c_int64, ctypes.c_int32]
            L.tmol_lkball_workspace_bytes.restype = ctypes.c_int64
            L.tmol_lkball_census.argtypes = [
                SP, BP, KP, _VP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64, _VP,
                _VP]
            _bound = True
    return L


def _np(t):
    return t.detach().cpu().numpy()


class LKBallTables:
    """Device tables of lk_ball for one block-type set and dtype, and the C
    structs pointing at them (block types: n_atoms, n_conn, conn_atoms)."""

    def __init__(self, n_atoms, n_conn, conn_atoms, n_occ, n_polar, occ_atoms, occ_params,
                 recipes_list, path_distance, global_params, watergen_params, tors, dtype,
                 device):
        T = int(n_atoms.numel())
        Amax = int(path_distance.shape[1])
        Omax = max(1, int(occ_atoms.shape[1]))
        if int(n_occ.max()) > MAX_OCC:
            raise ops.LJLKLibraryError("lk_ball: more than %d heavy atoms in a block type"
                                       % MAX_OCC)
        Rmax = max([r.shape[0] for r in recipes_list] + [1])
        rec = numpy.zeros((T, Rmax, RECIPE_WIDTH), dtype=numpy.int32)
        n_rec = numpy.zeros(T, dtype=numpy.int32)
        for i, r in enumerate(recipes_list):
            rec[i, : r.shape[0]] = r
            n_rec[i] = r.shape[0]
        tors_t = numpy.zeros((3, 4), dtype=numpy.float64)
        for k, v in enumerate(tors):
            v = numpy.asarray(v, dtype=numpy.float64).reshape(-1)
            if v.size > 4:
                raise ops.LJLKLibraryError("lk_ball: more than 4 water torsions")
            tors_t[k, : v.size] = v
        f = lambda x: torch.as_tensor(x).to(device=device, dtype=dtype).contiguous()  # noqa
        i = lambda x: ops._i32(torch.as_tensor(x), device)  # noqa: E731
        self.t = dict(
            n_atoms=i(n_atoms), n_conn=i(n_conn), conn_atoms=i(conn_atoms), n_occ=i(n_occ),
            n_polar=i(n_polar), occ_atoms=i(occ_atoms), occ_params=f(occ_params),
            n_recipes=i(n_rec), recipes=i(rec), path_distance=i(path_distance),
            global_params=f(global_params).reshape(-1), watergen_params=f(watergen_params)
            .reshape(-1), water_tors=f(tors_t),
        )
        b = ops.CBlockTypes()
        b.n_block_types, b.max_n_atoms = T, Amax
        b.max_n_tiles = (Amax - 1) // 32 + 1
        b.max_n_conn = int(conn_atoms.shape[1]) if conn_atoms.dim() == 2 else 0
        b.n_atom_types = 0
        b.dtype = ops._dtype_code(dtype)
        for k in ("n_atoms", "n_conn"):
            setattr(b, k, ops._p(self.t[k]))
        b.conn_atoms = ops._p(self.t["conn_atoms"])
        self.cbt = b
        k = CLKBallTables()
        k.max_n_occ, k.max_n_recipes = Omax, int(Rmax)
        k.max_n_polar = int(numpy.asarray(n_polar).max()) if T else 0
        if k.max_n_polar > MAX_POLAR:
            raise ops.LJLKLibraryError("lk_ball: more than %d polar atoms in a block type"
                                       % MAX_POLAR)
        for name in ("n_occ", "n_polar", "occ_atoms", "n_recipes", "recipes", "path_distance"):
            setattr(k, name, ctypes.cast(self.t[name].data_ptr(), _P32))
        for name in ("occ_params", "global_params", "watergen_params", "water_tors"):
            setattr(k, name, self.t[name].data_ptr())
        self.ck = k
        self.max_dis = float(numpy.float32(_np(torch.as_tensor(global_params)).reshape(-1)[4]))

    @classmethod
    def from_term(cls, term, pbt, dtype, device):
        """Tables of the block types of a PackedBlockTypes, from the term
        (parameters as the float32 values the reference's op receives)."""
        bts = [term.block_type(rt) for rt in pbt.active_block_types]
        T = len(bts)
        Omax = max(1, max(len(b.occ_list) for b in bts))
        occ = numpy.full((T, Omax), -1, dtype=numpy.int32)
        prm = numpy.zeros((T, Omax, 9), dtype=numpy.float32)
        n_occ = numpy.zeros(T, dtype=numpy.int32)
        n_pol = numpy.zeros(T, dtype=numpy.int32)
        tp = term.lkb_type_params.numpy()
        for i, (b, rt) in enumerate(zip(bts, pbt.active_block_types)):
            n_occ[i], n_pol[i] = len(b.occ_list), b.n_polar
            occ[i, : len(b.occ_list)] = b.occ_list
            prm[i, : len(b.occ_list)] = tp[rt.atom_types[b.occ_list]]
        pd = term.packed_tables(pbt)["path_distance"]
        return cls(pbt.n_atoms, pbt.n_conn, pbt.conn_atom, n_occ, n_pol, occ, prm,
                   [b.recipes for b in bts], pd, term.global_params, term.watergen_params,
                   tuple(_np(t) for t in term.water_tors()), dtype, device)


def lkb_workspace(cstack, tables, capacity, dcode, device):
    """A workspace for lk_ball's neighbor pass and score."""
    n = _lkb_lib().tmol_lkball_workspace_bytes(ctypes.byref(cstack), ctypes.byref(tables.cbt),
                                               ctypes.byref(tables.ck), capacity, dcode)
    if n < 0:
        raise ops.LJLKLibraryError("lk_ball workspace size query failed")
    return ops.Workspace(cstack, tables.cbt, capacity, dcode, device, nbytes=int(n))


# ---------------------------------------------------------------- op-argument conversion

_conv_cache = {}


def _key(*ts):
    """Content key of the tables (small block-type tensors): a freed table's
    address can be reused by different tables."""
    h = hashlib.sha1()
    for t in ts:
        a = numpy.ascontiguousarray(_np(t))
        h.update(str((a.shape, a.dtype.str)).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def _recipes_from_op(n_atoms, all_bonds, ranges, t_ndonH, t_nacc, t_donH, t_donH_hvy, t_which,
                     t_acc, t_hyb, t_nH, is_h, n_tors):
    """Water recipes per block type from gen_pose_waters's tile tables
    (hbond_dependent_term.py:117-264 layout) and the all-bonds table."""
    key = ("rec", _key(n_atoms, all_bonds, ranges, t_ndonH, t_nacc, t_donH, t_donH_hvy, t_which,
                       t_acc, t_hyb, t_nH, is_h), tuple(n_tors))
    if key in _conv_cache:
        return _conv_cache[key]
    na, ab, rg = _np(n_atoms), _np(all_bonds), _np(ranges)
    nd, nac = _np(t_ndonH), _np(t_nacc)
    dH, dHv, wh = _np(t_donH), _np(t_donH_hvy), _np(t_which)
    ac, hy, nH = _np(t_acc), _np(t_hyb), _np(t_nH)
    ih = _np(is_h)
    out = []
    for bt in range(na.shape[0]):
        donH, D, which, acc, hyb, nHa = [], [], [], [], {}, {}
        for t in range(nd.shape[1]):
            for k in range(int(nd[bt, t])):
                donH.append(t * 32 + int(dH[bt, t, k]))
                D.append(int(dHv[bt, t, k]))
                which.append(int(wh[bt, t, k]))
            for k in range(int(nac[bt, t])):
                a = t * 32 + int(ac[bt, t, k])
                acc.append(a)
                hyb[a] = int(hy[bt, t, k])
                nHa[a] = int(nH[bt, t, k])
        out.append(water_recipes(donH, D, which, acc, hyb, nHa, ab[bt], rg[bt], ih[bt], n_tors,
                                 "block type %d" % bt))
    _conv_cache[key] = out
    return out


def _occluders_from_op(n_atoms, t_npol, t_nocc, t_poi, t_params):
    """Occluder lists (polar first over the whole block) and their
    parameters from lk_ball_pose_score's tile tables
    (lk_ball_energy_term.py:69-198 layout)."""
    key = ("occ", _key(n_atoms, t_npol, t_nocc, t_poi, t_params))
    if key in _conv_cache:
        return _conv_cache[key]
    na = _np(n_atoms)
    npol, nocc, poi = _np(t_npol), _np(t_nocc), _np(t_poi)
    prm = _np(t_params.float())
    T = na.shape[0]
    lists = []
    for bt in range(T):
        pol, apo = [], []
        for t in range(npol.shape[1]):
            for k in range(int(nocc[bt, t])):
                entry = (t * 32 + int(poi[bt, t, k]), prm[bt, t, k])
                (pol if k < int(npol[bt, t]) else apo).append(entry)
        lists.append((pol, apo))
    Omax = max(1, max(len(p) + len(a) for p, a in lists))
    occ = numpy.full((T, Omax), -1, dtype=numpy.int32)
    params = numpy.zeros((T, Omax, 9), dtype=numpy.float64)
    n_occ = numpy.zeros(T, dtype=numpy.int32)
    n_pol = numpy.zeros(T, dtype=numpy.int32)
    for bt, (pol, apo) in enumerate(lists):
        for k, (a, p) in enumerate(pol + apo):
            occ[bt, k] = a
            params[bt, k] = p
        n_occ[bt], n_pol[bt] = len(pol) + len(apo), len(pol)
    res = (n_occ, n_pol, occ, params)
    _conv_cache[key] = res
    return res


def _padded_cstack(rco, bt_for_rot, pia, mbs, ibs, P, B, NA, dev):
    keep = dict(
        block_atom_offset=ops._i32(rco, dev),
        block_type=ops._i32(bt_for_rot, dev),
        atom_pose=ops._i32(pia, dev),
    )
    if mbs is not None:
        keep["min_bondsep"] = ops._i32(mbs, dev)
        keep["inter_block_bondsep"] = ops._i32(ibs, dev)
    s = ops.CStack()
    s.layout = ops.PADDED
    s.n_poses, s.max_n_blocks = P, B
    s.n_blocks_total = P * B
    s.n_atoms_total = NA
    for k, v in keep.items():
        setattr(s, k, ops._p(v))
    return s, keep


class _WaterGenCall:
    def __init__(self, args):
        (coords, rco, pia, frfb, _frbt, _bifr, _pifr, btifr, _nrp, _rop, _nrb, _rob, _mx,
         irc, bt_n_atoms, bt_n_conn, bt_conn_atoms, _nab, all_bonds, ranges, t_ndonH, t_nacc,
         t_donH, t_donHhvy, t_which, t_acc, t_hyb, t_nH, is_h, wg, sp2, sp3, ring) = args
        ops._require_device(coords)
        dev = coords.device
        P, B = frfb.shape
        NA = coords.shape[0]
        if coords.dim() != 2 or coords.shape[1] != 3:
            raise ops.LJLKLibraryError("rot_coords must be [n_atoms, 3]")
        if rco.numel() != P * B or btifr.numel() != P * B:
            raise ops.LJLKLibraryError("rot index tensors must have P * B entries")
        n_tors = (sp2.numel(), sp3.numel(), ring.numel())
        rec = _recipes_from_op(bt_n_atoms, all_bonds, ranges, t_ndonH, t_nacc, t_donH,
                               t_donHhvy, t_which, t_acc, t_hyb, t_nH, is_h, n_tors)
        T = int(bt_n_atoms.numel())
        Amax = int(is_h.shape[1])
        z = numpy.zeros(T, dtype=numpy.int32)
        self.tables = LKBallTables(
            bt_n_atoms, bt_n_conn, bt_conn_atoms, z, z, numpy.full((T, 1), -1, numpy.int32),
            numpy.zeros((T, 1, 9)), rec, torch.full((T, Amax, Amax), 6, dtype=torch.int32),
            torch.tensor([0, 0, 0, 0, 6.0]), wg, (_np(sp2), _np(sp3), _np(ring)),
            coords.dtype, dev)
        self.cstack, self.keep = _padded_cstack(rco, btifr, pia, None, None, int(P), int(B),
                                                NA, dev)
        self.coords = coords.contiguous()
        self.dcode = ops._dtype_code(coords.dtype)
        self.device = dev

    def refs(self):
        return ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ctypes.byref(
            self.tables.ck)

    def waters(self):
        W = torch.empty((self.coords.shape[0], 4, 3), dtype=self.coords.dtype,
                        device=self.device)
        ops._check(_lkb_lib().tmol_lkball_waters(
            *self.refs(), ops._p(self.coords), self.dcode, ops._p(W),
            ops._stream(self.device)), "lk_ball waters")
        return W

    def backward(self, dW):
        g = torch.zeros_like(self.coords)
        ops._check(_lkb_lib().tmol_lkball_waters_backward(
            *self.refs(), ops._p(self.coords), self.dcode, ops._p(dW.contiguous()), ops._p(g),
            ops._stream(self.device)), "lk_ball waters backward")
        return g


class PoseWaterGenFunction(torch.autograd.Function):
    """gen_pose_waters forward / backward (PoseWaterGen,
    compiled.ops.cpp:22-281)."""

    @staticmethod
    def forward(ctx, *args):
        call = _WaterGenCall(args)
        ctx.call = call
        ctx.n_args = len(args)
        return call.waters()

    @staticmethod
    def backward(ctx, dW):
        if dW is None:
            return (None,) * ctx.n_args
        return (ctx.call.backward(dW.to(ctx.call.coords.dtype)),) + (None,) * (ctx.n_args - 1)


def gen_pose_waters(*args):
    """Drop-in for ``torch.ops.tmol_lk_ball.gen_pose_waters`` (33 arguments)."""
    if len(args) != 33:
        raise TypeError("gen_pose_waters takes 33 arguments (%d given)" % len(args))
    return PoseWaterGenFunction.apply(*args)


class _ScoreCall:
    def __init__(self, args):
        (coords, rco, pia, frfb, _frbt, _bifr, _pifr, btifr, _nrp, _rop, _nrb, _rob, _mx, _irc,
         mbs, ibs, bt_n_atoms, bt_n_conn, bt_conn_atoms, t_npol, t_nocc, t_poi, t_params, pd, gp,
         max_dis, waters, block_pair) = args
        ops._require_device(coords)
        dev = coords.device
        P, B = frfb.shape
        self.P, self.B = int(P), int(B)
        NA = coords.shape[0]
        if tuple(mbs.shape) != (P, B, B):
            raise ops.LJLKLibraryError("pose_stack_min_bond_separation must be [P, B, B]")
        if tuple(waters.shape) != (NA, 4, 3):
            raise ops.LJLKLibraryError("water_coords must be [n_atoms, 4, 3]")
        n_occ, n_pol, occ, params = _occluders_from_op(bt_n_atoms, t_npol, t_nocc, t_poi,
                                                       t_params)
        T = int(bt_n_atoms.numel())
        self.tables = LKBallTables(
            bt_n_atoms, bt_n_conn, bt_conn_atoms, n_occ, n_pol, occ, params,
            [numpy.zeros((0, RECIPE_WIDTH), numpy.int32)] * T, pd, gp,
            torch.tensor([0.0, 0, 0, 0]), ([], [], []), coords.dtype, dev)
        self.cstack, self.keep = _padded_cstack(rco, btifr, pia, mbs, ibs, self.P, self.B, NA,
                                                dev)
        self.coords = coords.contiguous()
        self.waters = waters.to(coords.dtype).contiguous()
        self.dtype = coords.dtype
        self.dcode = ops._dtype_code(coords.dtype)
        self.device = dev
        self.max_dis = float(max_dis)
        self.block_pair = bool(block_pair)
        pbs = torch.arange(0, P + 1, dtype=torch.int32) * B
        self.tiles = ops.neighbor_tiles(pbs, dev)
        self.ws = lkb_workspace(self.cstack, self.tables, ops.padded_row_capacity(self.P, self.B),
                                self.dcode, dev)

    def refs(self):
        return ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ctypes.byref(
            self.tables.ck)

    def neighbors(self, dense):
        ops._check(_lkb_lib().tmol_lkball_neighbors(
            ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ops._p(self.coords),
            self.dcode, self.max_dis, ops._p(self.tiles), self.tiles.shape[0], ops._p(self.ws.buf),
            self.ws.nbytes, self.ws.capacity, ops._p(dense), ops._stream(self.device)),
            "lk_ball neighbors")

    def score(self, flags, weights=None, energies=None, dxyz=None, dwat=None):
        ops._check(_lkb_lib().tmol_lkball_score(
            *self.refs(), ops._p(self.coords), ops._p(self.waters), self.dcode,
            ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, flags, ops._p(weights),
            ops._p(energies), ops._p(dxyz), ops._p(dwat), 0, ops._stream(self.device)),
            "lk_ball score")


class LKBallPoseScoreFunction(torch.autograd.Function):
    """lk_ball_pose_score forward / backward (LKBallPoseScoreOp,
    compiled.ops.cpp:356-838): the backward re-walks the neighbor list with
    the incoming gradient as per-pose (or per-block-pair) weights and
    returns gradients for the coordinates and for the waters."""

    @staticmethod
    def forward(ctx, *args):
        call = _ScoreCall(args)
        P, B = call.P, call.B
        dev = call.device
        nbr = torch.empty((P, B, B), dtype=torch.int32, device=dev)
        call.neighbors(nbr)
        shape = (4, P, B, B) if call.block_pair else (4, P)
        score = torch.empty(shape, dtype=call.dtype, device=dev)
        call.score(ops.BLOCK_PAIR if call.block_pair else 0, energies=score)
        ctx.call = call
        ctx.n_args = len(args)
        ctx.mark_non_differentiable(nbr)
        return score, nbr

    @staticmethod
    def backward(ctx, grad_score, grad_nbr):
        if grad_score is None:
            return (None,) * ctx.n_args
        call = ctx.call
        w = grad_score.to(call.dtype).contiguous()
        dxyz = torch.empty_like(call.coords)
        dwat = torch.empty_like(call.waters)
        flags = ops.WGRAD | ops.NO_ENERGY | (ops.BLOCK_PAIR if call.block_pair else 0)
        call.score(flags, weights=w, dxyz=dxyz, dwat=dwat)
        out = [None] * ctx.n_args
        out[0] = dxyz
        out[26] = dwat
        return tuple(out)


def lk_ball_pose_score(*args):
    """Drop-in for ``torch.ops.tmol_lk_ball.lk_ball_pose_score`` (28
    arguments); returns ``[score, block_neighbors]``."""
    if len(args) != 28:
        raise TypeError("lk_ball_pose_score takes 28 arguments (%d given)" % len(args))
    score, nbr = LKBallPoseScoreFunction.apply(*args)
    return [score, nbr]


# ---------------------------------------------------------------- ragged scorer


class LKBallRaggedScorer:
    """lk_ball over a ``RaggedPoseStack``: neighbor pass (lk_ball's max_dis)
    into its own work list, waters, then the persistent row kernel.
    ``energy_and_grad(coords, weights)`` gives energies [4, P] and the
    gradient of sum_t sum_p weights[t, p] E_t,p [NA, 3] (waters' chain rule
    included)."""

    def __init__(self, stack, term, device="cuda", dtype=torch.float32,
                 capacity: Optional[int] = None):
        from .scorer import ragged_cstack

        dev = torch.device(device)
        self.device, self.dtype = dev, dtype
        self.dcode = ops._dtype_code(dtype)
        self.stack = stack
        self.P, self.NA = stack.n_poses, stack.n_atoms_total
        self.cstack, self.keep = ragged_cstack(stack, dev)
        self.tables = LKBallTables.from_term(term, stack.packed_block_types, dtype, dev)
        self.max_dis = self.tables.max_dis
        self.tiles = ops.neighbor_tiles(stack.pose_block_start, dev)
        max_nb = int(stack.n_blocks_per_pose.max())
        self.capacity = int(capacity or max_nb)
        self.ws = lkb_workspace(self.cstack, self.tables, self.capacity, self.dcode, dev)
        self.waters = torch.empty((self.NA, 4, 3), dtype=dtype, device=dev)
        self.energies = torch.zeros((4, self.P), dtype=dtype, device=dev)
        self.dxyz = torch.zeros((self.NA, 3), dtype=dtype, device=dev)
        self.dwat = torch.zeros((self.NA, 4, 3), dtype=dtype, device=dev)
        self._checked = self.capacity >= max_nb

    def refs(self):
        return ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ctypes.byref(
            self.tables.ck)

    def neighbors(self, coords):
        ops._check(_lkb_lib().tmol_lkball_neighbors(
            ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ops._p(coords), self.dcode,
            self.max_dis, ops._p(self.tiles), self.tiles.shape[0], ops._p(self.ws.buf),
            self.ws.nbytes, self.ws.capacity, None, ops._stream(self.device)),
            "lk_ball neighbors")

    def gen_waters(self, coords):
        ops._check(_lkb_lib().tmol_lkball_waters(
            *self.refs(), ops._p(coords), self.dcode, ops._p(self.waters),
            ops._stream(self.device)), "lk_ball waters")
        return self.waters

    def score(self, coords, flags, weights=None, energies=None, dxyz=None, dwat=None,
              n_wave_groups=0):
        ops._check(_lkb_lib().tmol_lkball_score(
            *self.refs(), ops._p(coords), ops._p(self.waters), self.dcode, ops._p(self.ws.buf),
            self.ws.nbytes, self.ws.capacity, flags, ops._p(weights), ops._p(energies),
            ops._p(dxyz), ops._p(dwat), int(n_wave_groups), ops._stream(self.device)),
            "lk_ball score")

    def waters_backward(self, coords, dwat, dxyz):
        ops._check(_lkb_lib().tmol_lkball_waters_backward(
            *self.refs(), ops._p(coords), self.dcode, ops._p(dwat), ops._p(dxyz),
            ops._stream(self.device)), "lk_ball waters backward")

    def census(self, coords) -> dict:
        """Pair counts of the current work list (after neighbors and
        gen_waters): candidates, separation <= 3, evaluated, with bridge."""
        out = torch.zeros(4, dtype=torch.int64, device=self.device)
        ops._check(_lkb_lib().tmol_lkball_census(
            *self.refs(), ops._p(coords), ops._p(self.waters), self.dcode, ops._p(self.ws.buf),
            self.ws.nbytes, self.ws.capacity, ops._p(out), ops._stream(self.device)),
            "lk_ball census")
        return dict(zip(("candidates", "cp0", "evaluated", "bridge"), out.cpu().tolist()))

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
            w = w.expand(4, self.P)
        elif w.dim() == 1:
            w = w[:, None].expand(4, self.P)
        return w.contiguous()

    def energies_only(self, coords):
        self.neighbors(coords)
        self.gen_waters(coords)
        self.score(coords, 0, energies=self.energies)
        return self.energies

    def energy_and_grad(self, coords, weights=None):
        w = self._weights(weights)
        self.neighbors(coords)
        self.gen_waters(coords)
        self.score(coords, ops.WGRAD, weights=w, energies=self.energies, dxyz=self.dxyz,
                   dwat=self.dwat)
        self.waters_backward(coords, self.dwat, self.dxyz)
        if not self._checked:
            self.check_capacity()
            self._checked = True
        return self.energies, self.dxyz
