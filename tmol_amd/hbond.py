"""hbond on the HIP library: device tables, the drop-in pose ops and a
ragged-batch scorer.

Drop-ins (same argument lists and outputs as the reference's ops, on device
tensors):

* ``gen_hbond_bases(*20 args)`` -> (derived_coords [NA, 3, 3],
  derived_atom_inds [NA, 3] int32), for ``tmol_hbond::gen_hbond_bases``
  (tmol/score/hbond/potentials/compiled.ops.cpp:930-996);
* ``hbond_pose_scores(*37 args)`` -> [score, block_neighbors], score [1, P]
  or [1, P, B, B], for ``tmol_hbond::hbond_pose_scores``
  (compiled.ops.cpp:28-375, 723-825), differentiable in the coordinates
  (gradients reach the base atoms through derived_atom_inds, as in the
  reference).

The reference's block-type tables (tile-ordered donor hydrogens and
acceptors with their types and hybridizations, the all-bonds table) are
converted once per table set into this library's form: per block type the
donor list {H, D, type}, the acceptor list {A, B, B0, type, hybridization}
and the source atom of each derived slot.

``HBondRaggedScorer`` scores a RaggedPoseStack: neighbor pass, bases,
energies and the weighted gradient.  Nothing here falls back to a CPU path:
without the HIP library every entry point raises.
"""

import ctypes
import threading
from typing import Optional

import numpy
import torch

from . import ops
from .lk_ball import _key, _np, _padded_cstack
from .score.hbond_energy_term import ACC_WIDTH, DON_WIDTH, NEIGHBOR_REACH, acceptor_bases

_VP = ctypes.c_void_p
_P32 = ctypes.POINTER(ctypes.c_int32)
MAX_LIST = 32


class CHBondTables(ctypes.Structure):
    _fields_ = [
        ("max_n_don", ctypes.c_int32),
        ("max_n_acc", ctypes.c_int32),
        ("n_don_types", ctypes.c_int32),
        ("n_acc_types", ctypes.c_int32),
        ("n_don", _P32),
        ("n_acc", _P32),
        ("don", _P32),
        ("acc", _P32),
        ("base_src", _P32),
        ("path_distance", _P32),
        ("pair_params", _VP),
        ("pair_polys", _VP),
        ("global_params", _VP),
    ]


_bound = False
_lock = threading.Lock()


def _hb_lib():
    global _bound
    L = ops.lib()
    with _lock:
        if not _bound:
            SP, BP, HP = (ctypes.POINTER(ops.CStack), ctypes.POINTER(ops.CBlockTypes),
                          ctypes.POINTER(CHBondTables))
            L.tmol_hbond_bases.argtypes = [SP, BP, HP, _VP, ctypes.c_int32, _VP, _VP, _VP]
            L.tmol_hbond_neighbors.argtypes = [
                SP, BP, _VP, ctypes.c_int32, ctypes.c_double, _VP, ctypes.c_int64, _VP,
                ctypes.c_int64, ctypes.c_int64, _VP, _VP]
            L.tmol_hbond_score.argtypes = [
                SP, BP, HP, _VP, _VP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64,
                ctypes.c_int32, _VP, _VP, _VP, ctypes.c_int32, _VP]
            L.tmol_hbond_census.argtypes = [
                SP, BP, HP, _VP, _VP, _VP, ctypes.c_int32, _VP, ctypes.c_int64, ctypes.c_int64,
                _VP, _VP]
            _bound = True
    return L


def _base_src(T, Amax, don_lists, acc_lists):
    """[T, Amax, 3] source atom of each derived slot: slot 0 the D of a
    polar H, slots 1, 2 the B, B0 of an acceptor; -1 missing base, -2 no
    role."""
    src = numpy.full((T, Amax, 3), -2, dtype=numpy.int32)
    for bt in range(T):
        for h, d, _t, _ in don_lists[bt]:
            src[bt, h, 0] = d
        for a, b, b0, _t, _hyb, _, _, _ in acc_lists[bt]:
            src[bt, a, 1] = b
            src[bt, a, 2] = b0
    return src


class HBondTables:
    """Device tables of hbond for one block-type set and dtype, and the C
    structs pointing at them (block types: n_atoms, n_conn, conn_atoms)."""

    def __init__(self, n_atoms, n_conn, conn_atoms, don_lists, acc_lists, path_distance,
                 pair_params, pair_polys, global_params, dtype, device):
        T = int(n_atoms.numel())
        Amax = int(path_distance.shape[1])
        nd = numpy.array([len(x) for x in don_lists], dtype=numpy.int32)
        na = numpy.array([len(x) for x in acc_lists], dtype=numpy.int32)
        if T and (int(nd.max()) > MAX_LIST or int(na.max()) > MAX_LIST):
            raise ops.LJLKLibraryError("hbond: more than %d donors or acceptors in a block type"
                                       % MAX_LIST)
        Dmax, Xmax = max(1, int(nd.max()) if T else 1), max(1, int(na.max()) if T else 1)
        don = numpy.zeros((T, Dmax, DON_WIDTH), dtype=numpy.int32)
        acc = numpy.zeros((T, Xmax, ACC_WIDTH), dtype=numpy.int32)
        for bt in range(T):
            if nd[bt]:
                don[bt, : nd[bt]] = don_lists[bt]
            if na[bt]:
                acc[bt, : na[bt]] = acc_lists[bt]
        src = _base_src(T, Amax, don_lists, acc_lists)
        f = lambda x: torch.as_tensor(x).to(device=device, dtype=dtype).contiguous()  # noqa
        i = lambda x: ops._i32(torch.as_tensor(x), device)  # noqa: E731
        pp = torch.as_tensor(pair_params)
        self.t = dict(
            n_atoms=i(n_atoms), n_conn=i(n_conn), conn_atoms=i(conn_atoms), n_don=i(nd),
            n_acc=i(na), don=i(don), acc=i(acc), base_src=i(src),
            path_distance=i(path_distance), pair_params=f(pp),
            pair_polys=torch.as_tensor(pair_polys).to(device=device, dtype=torch.float64)
            .contiguous(), global_params=f(global_params).reshape(-1),
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
        h = CHBondTables()
        h.max_n_don, h.max_n_acc = Dmax, Xmax
        h.n_don_types, h.n_acc_types = int(pp.shape[0]), int(pp.shape[1])
        for name in ("n_don", "n_acc", "don", "acc", "base_src", "path_distance"):
            setattr(h, name, ctypes.cast(self.t[name].data_ptr(), _P32))
        for name in ("pair_params", "pair_polys", "global_params"):
            setattr(h, name, self.t[name].data_ptr())
        self.ch = h
        self.max_ha_dis = float(_np(torch.as_tensor(global_params)).reshape(-1)[5])

    @classmethod
    def from_term(cls, term, pbt, dtype, device):
        """Tables of the block types of a PackedBlockTypes, from the term
        (parameters as the values the reference's op receives)."""
        bts = [term.block_type(rt) for rt in pbt.active_block_types]
        pd = term.packed_tables(pbt)["path_distance"]
        return cls(pbt.n_atoms, pbt.n_conn, pbt.conn_atom, [b.don for b in bts],
                   [b.acc for b in bts], pd, term.pair_param_table, term.pair_poly_table,
                   term.global_param_table, dtype, device)


def hbond_workspace(cstack, tables, capacity, dcode, device):
    """A workspace for hbond's neighbor pass and score (the LJ+LK layout)."""
    return ops.Workspace(cstack, tables.cbt, capacity, dcode, device)


# ---------------------------------------------------------------- op-argument conversion

_conv_cache = {}


def _lists_from_op(n_atoms, all_bonds, ranges, t_ndonH, t_nacc, t_donH, t_acc, t_hyb, is_h,
                   t_dtype=None, t_atype=None):
    """Per block type the donor list {H, D, type, 0} and the acceptor list
    {A, B, B0, type, hybridization, 0, 0, 0} from the ops' tile tables
    (hbond_dependent_term.py:117-264 layout) and the all-bonds table; D is
    the last heavy partner of H, B / B0 by the reference's walk
    (identification.hh:26-198)."""
    key = ("hb", _key(n_atoms, all_bonds, ranges, t_ndonH, t_nacc, t_donH, t_acc, t_hyb, is_h,
                      *([t_dtype, t_atype] if t_dtype is not None else [])))
    if key in _conv_cache:
        return _conv_cache[key]
    na, ab, rg = _np(n_atoms), _np(all_bonds), _np(ranges)
    nd, nac = _np(t_ndonH), _np(t_nacc)
    dH, ac, hy = _np(t_donH), _np(t_acc), _np(t_hyb)
    ih = _np(is_h)
    dt = _np(t_dtype) if t_dtype is not None else None
    at = _np(t_atype) if t_atype is not None else None
    dons, accs = [], []
    for bt in range(na.shape[0]):
        d_rows, a_rows = [], []
        for t in range(nd.shape[1]):
            for k in range(int(nd[bt, t])):
                h = t * 32 + int(dH[bt, t, k])
                D = -1
                for e in range(rg[bt, h, 0], rg[bt, h, 1]):
                    x = int(ab[bt, e, 1])
                    if x < 0:
                        raise NotImplementedError(
                            "hbond: polar hydrogen %d of block type %d is bonded across "
                            "blocks" % (h, bt))
                    if not ih[bt, x]:
                        D = x
                d_rows.append([h, D, int(dt[bt, t, k]) if dt is not None else 0, 0])
            for k in range(int(nac[bt, t])):
                a = t * 32 + int(ac[bt, t, k])
                hyb = int(hy[bt, t, k])
                B, B0 = acceptor_bases(a, hyb, ab[bt], rg[bt], ih[bt], "block type %d" % bt)
                a_rows.append([a, B, B0, int(at[bt, t, k]) if at is not None else 0, hyb, 0, 0,
                               0])
        dons.append(numpy.array(d_rows, dtype=numpy.int32).reshape(-1, DON_WIDTH))
        accs.append(numpy.array(a_rows, dtype=numpy.int32).reshape(-1, ACC_WIDTH))
    _conv_cache[key] = (dons, accs)
    return dons, accs


_DUMMY_POLYS = numpy.zeros((1, 1, 48))


class _BasesCall:
    def __init__(self, args):
        (coords, rco, frfb, _frbt, _bifr, _pifr, btifr, _irc, bt_n_atoms, bt_n_conn,
         bt_conn_atoms, _nab, all_bonds, ranges, t_ndonH, t_nacc, t_donH, t_acc, t_hyb,
         is_h) = args
        ops._require_device(coords)
        dev = coords.device
        P, B = frfb.shape
        NA = coords.shape[0]
        if coords.dim() != 2 or coords.shape[1] != 3:
            raise ops.LJLKLibraryError("rot_coords must be [n_atoms, 3]")
        if rco.numel() != P * B or btifr.numel() != P * B:
            raise ops.LJLKLibraryError("rot index tensors must have P * B entries")
        dons, accs = _lists_from_op(bt_n_atoms, all_bonds, ranges, t_ndonH, t_nacc, t_donH,
                                    t_acc, t_hyb, is_h)
        T = int(bt_n_atoms.numel())
        Amax = int(is_h.shape[1])
        self.tables = HBondTables(
            bt_n_atoms, bt_n_conn, bt_conn_atoms, dons, accs,
            torch.full((T, Amax, Amax), 6, dtype=torch.int32), numpy.zeros((1, 1, 3)),
            _DUMMY_POLYS, torch.zeros(6), coords.dtype, dev)
        self.cstack, self.keep = _padded_cstack(rco, btifr, torch.zeros(NA, dtype=torch.int32),
                                                None, None, int(P), int(B), NA, dev)
        self.coords = coords.contiguous()
        self.dcode = ops._dtype_code(coords.dtype)
        self.device = dev

    def bases(self):
        NA = self.coords.shape[0]
        D = torch.empty((NA, 3, 3), dtype=self.coords.dtype, device=self.device)
        I = torch.empty((NA, 3), dtype=torch.int32, device=self.device)
        ops._check(_hb_lib().tmol_hbond_bases(
            ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt),
            ctypes.byref(self.tables.ch), ops._p(self.coords), self.dcode, ops._p(D), ops._p(I),
            ops._stream(self.device)), "hbond bases")
        return D, I


def gen_hbond_bases(*args):
    """Drop-in for ``torch.ops.tmol_hbond.gen_hbond_bases`` (20 arguments);
    returns (derived_coords, derived_atom_inds)."""
    if len(args) != 20:
        raise TypeError("gen_hbond_bases takes 20 arguments (%d given)" % len(args))
    with torch.no_grad():
        return _BasesCall(args).bases()


class _ScoreCall:
    def __init__(self, args):
        (coords, rco, pia, frfb, _frbt, _bifr, _pifr, btifr, _nrp, _rop, _nrb, _rob, _mx, _irc,
         mbs, ibs, bt_n_atoms, bt_n_conn, bt_conn_atoms, _nab, all_bonds, ranges, pd, t_ndonH,
         t_nacc, t_donH, t_acc, t_dtype, t_atype, t_hyb, is_h, pair_params, pair_polys, gp,
         derived, dinds, block_pair) = args
        ops._require_device(coords)
        dev = coords.device
        P, B = frfb.shape
        self.P, self.B = int(P), int(B)
        NA = coords.shape[0]
        if tuple(mbs.shape) != (P, B, B):
            raise ops.LJLKLibraryError("pose_stack_min_bond_separation must be [P, B, B]")
        if tuple(derived.shape) != (NA, 3, 3) or tuple(dinds.shape) != (NA, 3):
            raise ops.LJLKLibraryError("derived coords / atom inds must be [n_atoms, 3(, 3)]")
        dons, accs = _lists_from_op(bt_n_atoms, all_bonds, ranges, t_ndonH, t_nacc, t_donH,
                                    t_acc, t_hyb, is_h, t_dtype, t_atype)
        self.tables = HBondTables(bt_n_atoms, bt_n_conn, bt_conn_atoms, dons, accs, pd,
                                  pair_params, pair_polys, gp, coords.dtype, dev)
        self.cstack, self.keep = _padded_cstack(rco, btifr, pia, mbs, ibs, self.P, self.B, NA,
                                                dev)
        self.coords = coords.contiguous()
        self.derived = derived.to(coords.dtype).contiguous()
        self.dinds = ops._i32(dinds, dev)
        self.dtype = coords.dtype
        self.dcode = ops._dtype_code(coords.dtype)
        self.device = dev
        self.block_pair = bool(block_pair)
        pbs = torch.arange(0, P + 1, dtype=torch.int32) * B
        self.tiles = ops.neighbor_tiles(pbs, dev)
        self.ws = hbond_workspace(self.cstack, self.tables,
                                  ops.padded_row_capacity(self.P, self.B), self.dcode, dev)

    def refs(self):
        return ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ctypes.byref(
            self.tables.ch)

    def neighbors(self, dense):
        ops._check(_hb_lib().tmol_hbond_neighbors(
            ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ops._p(self.coords),
            self.dcode, NEIGHBOR_REACH, ops._p(self.tiles), self.tiles.shape[0],
            ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, ops._p(dense),
            ops._stream(self.device)), "hbond neighbors")

    def score(self, flags, weights=None, energies=None, dxyz=None):
        ops._check(_hb_lib().tmol_hbond_score(
            *self.refs(), ops._p(self.coords), ops._p(self.derived), ops._p(self.dinds),
            self.dcode, ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, flags,
            ops._p(weights), ops._p(energies), ops._p(dxyz), 0, ops._stream(self.device)),
            "hbond score")


class HBondPoseScoreFunction(torch.autograd.Function):
    """hbond_pose_scores forward / backward (HBondPoseScoresOp,
    compiled.ops.cpp:28-375): the backward re-walks the neighbor list with
    the incoming gradient as per-pose (or per-block-pair) weights."""

    @staticmethod
    def forward(ctx, *args):
        call = _ScoreCall(args)
        P, B = call.P, call.B
        dev = call.device
        nbr = torch.empty((P, B, B), dtype=torch.int32, device=dev)
        call.neighbors(nbr)
        shape = (1, P, B, B) if call.block_pair else (1, P)
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
        flags = ops.WGRAD | ops.NO_ENERGY | (ops.BLOCK_PAIR if call.block_pair else 0)
        call.score(flags, weights=w, dxyz=dxyz)
        out = [None] * ctx.n_args
        out[0] = dxyz
        return tuple(out)


def hbond_pose_scores(*args):
    """Drop-in for ``torch.ops.tmol_hbond.hbond_pose_scores`` (37
    arguments); returns ``[score, block_neighbors]``."""
    if len(args) != 37:
        raise TypeError("hbond_pose_scores takes 37 arguments (%d given)" % len(args))
    score, nbr = HBondPoseScoreFunction.apply(*args)
    return [score, nbr]


# ---------------------------------------------------------------- ragged scorer


class HBondRaggedScorer:
    """hbond over a ``RaggedPoseStack``: neighbor pass (5.5 A) into its own
    half-triangle work list, bases, then the persistent row kernel.
    ``energy_and_grad(coords, weights)`` gives energies [P] and the gradient
    of sum_p weights[p] E_p [NA, 3]."""

    def __init__(self, stack, term, device="cuda", dtype=torch.float32,
                 capacity: Optional[int] = None):
        from .scorer import ragged_cstack

        dev = torch.device(device)
        self.device, self.dtype = dev, dtype
        self.dcode = ops._dtype_code(dtype)
        self.stack = stack
        self.P, self.NA = stack.n_poses, stack.n_atoms_total
        self.cstack, self.keep = ragged_cstack(stack, dev)
        self.tables = HBondTables.from_term(term, stack.packed_block_types, dtype, dev)
        self.tiles = ops.neighbor_tiles(stack.pose_block_start, dev)
        max_nb = int(stack.n_blocks_per_pose.max())
        self.capacity = int(capacity or max_nb)
        self.ws = hbond_workspace(self.cstack, self.tables, self.capacity, self.dcode, dev)
        self.derived = torch.empty((self.NA, 3, 3), dtype=dtype, device=dev)
        self.dinds = torch.empty((self.NA, 3), dtype=torch.int32, device=dev)
        self.energies = torch.zeros(self.P, dtype=dtype, device=dev)
        self.dxyz = torch.zeros((self.NA, 3), dtype=dtype, device=dev)
        self._checked = self.capacity >= max_nb

    def refs(self):
        return ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ctypes.byref(
            self.tables.ch)

    def neighbors(self, coords):
        ops._check(_hb_lib().tmol_hbond_neighbors(
            ctypes.byref(self.cstack), ctypes.byref(self.tables.cbt), ops._p(coords), self.dcode,
            NEIGHBOR_REACH, ops._p(self.tiles), self.tiles.shape[0], ops._p(self.ws.buf),
            self.ws.nbytes, self.ws.capacity, None, ops._stream(self.device)),
            "hbond neighbors")

    def bases(self, coords):
        ops._check(_hb_lib().tmol_hbond_bases(
            *self.refs(), ops._p(coords), self.dcode, ops._p(self.derived), ops._p(self.dinds),
            ops._stream(self.device)), "hbond bases")
        return self.derived, self.dinds

    def score(self, coords, flags, weights=None, energies=None, dxyz=None, n_wave_groups=0):
        ops._check(_hb_lib().tmol_hbond_score(
            *self.refs(), ops._p(coords), ops._p(self.derived), ops._p(self.dinds), self.dcode,
            ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, flags, ops._p(weights),
            ops._p(energies), ops._p(dxyz), int(n_wave_groups), ops._stream(self.device)),
            "hbond score")

    def census(self, coords) -> dict:
        """Pair counts of the current work list (after neighbors and bases):
        candidates, below separation 5, evaluated."""
        out = torch.zeros(3, dtype=torch.int64, device=self.device)
        ops._check(_hb_lib().tmol_hbond_census(
            *self.refs(), ops._p(coords), ops._p(self.derived), ops._p(self.dinds), self.dcode,
            ops._p(self.ws.buf), self.ws.nbytes, self.ws.capacity, ops._p(out),
            ops._stream(self.device)), "hbond census")
        return dict(zip(("candidates", "cp_excluded", "evaluated"), out.cpu().tolist()))

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

    def energies_only(self, coords):
        self.neighbors(coords)
        self.bases(coords)
        self.score(coords, 0, energies=self.energies)
        return self.energies

    def energy_and_grad(self, coords, weights=None):
        w = self._weights(weights)
        self.neighbors(coords)
        self.bases(coords)
        self.score(coords, ops.WGRAD, weights=w, energies=self.energies, dxyz=self.dxyz)
        if not self._checked:
            self.check_capacity()
            self._checked = True
        return self.energies, self.dxyz
