"""ctypes driver for the C oracle (liboracle_ljlk.so) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, and only as the checker.  ``oracle_pose_scores`` takes the same
27 arguments as ``ljlk_pose_scores`` (tmol/score/ljlk/potentials/
compiled.ops.cpp:533-597) and runs the serial CPU restatement.
"""

import ctypes
import os
import subprocess

import numpy
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_ljlk.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "tmol_ljlk_ref.so")

_P = ctypes.POINTER(ctypes.c_int32)


class OracleStack(ctypes.Structure):
    _fields_ = [
        ("n_poses", ctypes.c_int32),
        ("max_n_blocks", ctypes.c_int32),
        ("n_block_types", ctypes.c_int32),
        ("max_n_block_atoms", ctypes.c_int32),
        ("max_n_tiles", ctypes.c_int32),
        ("max_n_conn", ctypes.c_int32),
        ("rot_coord_offset", _P),
        ("block_type", _P),
        ("min_bondsep", _P),
        ("inter_block_bondsep", _P),
        ("bt_n_atoms", _P),
        ("bt_n_heavy_in_tile", _P),
        ("bt_heavy_in_tile", _P),
        ("bt_atom_types", _P),
        ("bt_n_conn", _P),
        ("bt_conn_atoms", _P),
        ("bt_path_distance", _P),
        ("bt_is_ligand", _P),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for suf, ct in (("_f", ctypes.c_float), ("_d", ctypes.c_double)):
            PR = ctypes.POINTER(ct)
            f = getattr(L, "oracle_block_spheres" + suf)
            f.argtypes = [PR, ctypes.c_int, _P, _P, _P, PR]
            f = getattr(L, "oracle_block_neighbors" + suf)
            f.argtypes = [ctypes.c_int, ctypes.c_int, _P, PR, ct, _P]
            f = getattr(L, "oracle_pose_scores" + suf)
            f.argtypes = [
                ctypes.POINTER(OracleStack), PR, ctypes.c_int64, PR, PR, _P,
                ctypes.c_int, ctypes.c_int, PR, PR,
            ]
            f = getattr(L, "oracle_lj_pair" + suf)
            f.argtypes = [ct, ctypes.c_int, PR, PR, PR, ctypes.c_int, PR]
            f = getattr(L, "oracle_lk_pair" + suf)
            f.argtypes = [ct, ctypes.c_int, PR, PR, PR, ctypes.c_int, PR]
            f = getattr(L, "oracle_rot_block_spheres" + suf)
            f.argtypes = [ctypes.c_int, ctypes.c_int, PR, _P, _P, PR]
            f = getattr(L, "oracle_rotamer_scores" + suf)
            f.argtypes = [
                ctypes.POINTER(OracleStack), PR, ctypes.c_int64, PR, PR, ctypes.c_int64, _P,
                _P, _P, _P, PR, PR, PR,
            ]
        for suf, ct in (("_f", ctypes.c_float), ("_d", ctypes.c_double)):
            PR = ctypes.POINTER(ct)
            f = getattr(L, "oracle_elec_pair" + suf)
            f.argtypes = [ct, ct, ct, ctypes.c_int, PR, ctypes.c_int, PR]
            f = getattr(L, "oracle_elec_scores" + suf)
            f.argtypes = [
                ctypes.POINTER(OracleStack), PR, ctypes.c_int64, PR, _P, PR, _P,
                ctypes.c_int, ctypes.c_int, PR, PR,
            ]
            f = getattr(L, "oracle_elec_rotamer_scores" + suf)
            f.argtypes = [
                ctypes.POINTER(OracleStack), PR, ctypes.c_int64, PR, _P, PR, ctypes.c_int64, _P,
                _P, _P, _P, PR, PR, PR,
            ]
        for suf, ct in (("_f", ctypes.c_float), ("_d", ctypes.c_double)):
            PR = ctypes.POINTER(ct)
            LS = ctypes.POINTER(OracleLKBStack)
            getattr(L, "oracle_lkb_waters" + suf).argtypes = [LS, PR, ctypes.c_int64, PR, PR,
                                                               PR, PR, PR]
            getattr(L, "oracle_lkb_waters_backward" + suf).argtypes = [LS, PR, PR, PR, PR, PR,
                                                                        PR, PR]
            getattr(L, "oracle_lkb_score" + suf).argtypes = [
                LS, PR, PR, ctypes.c_int64, PR, PR, _P, ctypes.c_int, PR, PR, PR, PR]
            getattr(L, "oracle_lkb_pair" + suf).argtypes = [PR, PR, PR, PR, ctypes.c_int, PR,
                                                             PR, PR, PR, PR]
            f = getattr(L, "oracle_lkb_fraction" + suf)
            f.argtypes = [PR, PR, ct, PR, PR]
            f.restype = ct
            f = getattr(L, "oracle_lkb_bridge" + suf)
            f.argtypes = [PR, PR, PR, PR, ct, PR, PR, PR, PR]
            f.restype = ct
            getattr(L, "oracle_lkb_build_water" + suf).argtypes = [
                ctypes.c_int, PR, PR, PR, ct, ct, ct, PR, PR, PR, PR, PR]
        for suf, ct in (("_f", ctypes.c_float), ("_d", ctypes.c_double)):
            PR = ctypes.POINTER(ct)
            PD = ctypes.POINTER(ctypes.c_double)
            HS = ctypes.POINTER(OracleHBStack)
            getattr(L, "oracle_hb_bases" + suf).argtypes = [HS, PR, ctypes.c_int64, PR, _P]
            getattr(L, "oracle_hb_score_pose" + suf).argtypes = [
                HS, PR, ctypes.c_int64, PR, _P, PR, PD, PR, _P, ctypes.c_int, PR, PR, PR]
            f = getattr(L, "oracle_hb_score" + suf)
            f.argtypes = [PR, PR, PR, PR, PR, PR, PD, PR, PR]
            f.restype = ct
        L.oracle_rot_dispatch.argtypes = [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P,
                                          ctypes.c_int64]
        L.oracle_rot_dispatch.restype = ctypes.c_int64
        L.oracle_inter_block_bondsep.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P]
        L.oracle_pair_sep.argtypes = [
            ctypes.POINTER(OracleStack), ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int,
        ]
        L.oracle_pair_sep.restype = ctypes.c_int
        L.oracle_block_pair_seps.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_void_p]
        L.oracle_block_pair_seps.restype = None
        _lib = L
    return _lib


def _np_i32(t):
    return numpy.ascontiguousarray(t.detach().cpu().numpy().astype(numpy.int32))


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class _Stack:
    """Keeps the numpy buffers alive for the struct's pointers."""

    def __init__(self, args):
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot,
         block_type_ind_for_rot, n_rots_for_pose, rot_offset_for_pose,
         n_rots_for_block, rot_offset_for_block, max_n_rots_per_pose,
         min_bondsep, inter_block_bondsep, bt_n_atoms, bt_n_heavy_in_tile,
         bt_heavy_in_tile, bt_atom_types, bt_n_conn, bt_conn_atoms,
         bt_path_distance, bt_is_ligand, type_params, global_params, max_dis,
         block_pair) = args
        P, B = first_rot_for_block.shape
        self.arrays = {
            "rot_coord_offset": _np_i32(rot_coord_offset),
            "block_type": _np_i32(block_type_ind_for_rot),
            "min_bondsep": _np_i32(min_bondsep),
            "inter_block_bondsep": _np_i32(inter_block_bondsep),
            "bt_n_atoms": _np_i32(bt_n_atoms),
            "bt_n_heavy_in_tile": _np_i32(bt_n_heavy_in_tile),
            "bt_heavy_in_tile": _np_i32(bt_heavy_in_tile),
            "bt_atom_types": _np_i32(bt_atom_types),
            "bt_n_conn": _np_i32(bt_n_conn),
            "bt_conn_atoms": _np_i32(bt_conn_atoms),
            "bt_path_distance": _np_i32(bt_path_distance),
            "bt_is_ligand": _np_i32(bt_is_ligand),
        }
        st = OracleStack()
        st.n_poses = P
        st.max_n_blocks = B
        st.n_block_types = bt_n_atoms.shape[0]
        st.max_n_block_atoms = bt_atom_types.shape[1]
        st.max_n_tiles = bt_n_heavy_in_tile.shape[1]
        st.max_n_conn = bt_conn_atoms.shape[1]
        for k, v in self.arrays.items():
            setattr(st, k, _ptr(v, ctypes.c_int32))
        self.struct = st
        self.P, self.B = P, B
        self.coords = coords
        self.type_params = type_params
        self.global_params = global_params
        self.max_dis = float(max_dis)
        self.block_pair = bool(block_pair)


def oracle_pose_scores(*args, require_grad=False, dtype=None):
    """Run the oracle on the op's 27 arguments.

    Returns (scores, block_neighbors, dV_dx): scores is (3, P) or (3, P, B, B),
    block_neighbors (P, B, B) int32, dV_dx (3, NA, 3) or None."""
    S = _Stack(args)
    if dtype is None:
        dtype = numpy.float64 if S.coords.dtype == torch.float64 else numpy.float32
    coords = numpy.ascontiguousarray(S.coords.detach().cpu().numpy().astype(dtype))
    return _scores(S, coords, require_grad, dtype)


def _scores(S, coords, require_grad, dtype):
    """The oracle on a prepared stack and host coordinates [NA, 3]."""
    L = lib()
    suf, ct = ("_d", ctypes.c_double) if dtype == numpy.float64 else ("_f", ctypes.c_float)
    tp = numpy.ascontiguousarray(S.type_params.detach().cpu().numpy().astype(dtype))
    gp = numpy.ascontiguousarray(S.global_params.detach().cpu().numpy().astype(dtype).reshape(-1))
    NA = coords.shape[0]
    PB = S.P * S.B
    spheres = numpy.zeros((PB, 4), dtype=dtype)
    getattr(L, "oracle_block_spheres" + suf)(
        _ptr(coords, ct), PB, _ptr(S.arrays["rot_coord_offset"], ctypes.c_int32),
        _ptr(S.arrays["block_type"], ctypes.c_int32),
        _ptr(S.arrays["bt_n_atoms"], ctypes.c_int32), _ptr(spheres, ct),
    )
    nbr = numpy.zeros((S.P, S.B, S.B), dtype=numpy.int32)
    # reach is the host scalar max_dis cast to the coordinate type
    getattr(L, "oracle_block_neighbors" + suf)(
        S.P, S.B, _ptr(S.arrays["block_type"], ctypes.c_int32), _ptr(spheres, ct),
        ct(dtype(S.max_dis)), _ptr(nbr, ctypes.c_int32),
    )
    if S.block_pair:
        out = numpy.zeros((3, S.P, S.B, S.B), dtype=dtype)
    else:
        out = numpy.zeros((3, S.P), dtype=dtype)
    dvdx = numpy.zeros((3, NA, 3), dtype=dtype) if require_grad else None
    getattr(L, "oracle_pose_scores" + suf)(
        ctypes.byref(S.struct), _ptr(coords, ct), NA, _ptr(tp, ct), _ptr(gp, ct),
        _ptr(nbr, ctypes.c_int32), int(S.block_pair), int(require_grad),
        _ptr(out, ct), _ptr(dvdx, ct) if dvdx is not None else None,
    )
    return out, nbr, dvdx, spheres


class _RotOracle:
    """Host arrays of one rotamer-op call (the 27 arguments of
    ``ljlk_rotamer_scores``) for the C oracle."""

    def __init__(self, args, dtype=None):
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot, block_type_ind_for_rot,
         n_rots_for_pose, rot_offset_for_pose, n_rots_for_block, rot_offset_for_block,
         max_n_rots_per_pose, min_bondsep, inter_block_bondsep, *rest) = args
        self.S = S = _Stack(args)
        if dtype is None:
            dtype = numpy.float64 if S.coords.dtype == torch.float64 else numpy.float32
        self.dtype = dtype
        self.suf, self.ct = ("_d", ctypes.c_double) if dtype == numpy.float64 else ("_f", ctypes.c_float)
        self.x = numpy.ascontiguousarray(S.coords.detach().cpu().numpy().astype(dtype))
        self.tp = numpy.ascontiguousarray(S.type_params.detach().cpu().numpy().astype(dtype))
        self.gp = numpy.ascontiguousarray(
            S.global_params.detach().cpu().numpy().astype(dtype).reshape(-1))
        self.rco = _np_i32(rot_coord_offset)
        self.btr = _np_i32(block_type_ind_for_rot)
        self.bir = _np_i32(block_ind_for_rot)
        self.nrb = _np_i32(n_rots_for_block)
        self.rob = _np_i32(rot_offset_for_block)
        self.frt = _np_i32(first_rot_block_type)

    def dispatch(self):
        """(indices [3, n] int32, rot_spheres [R, 4], block_spheres [P, B, 4],
        block_neighbors [P, B, B])."""
        L, S, ct, suf = lib(), self.S, self.ct, self.suf
        I = lambda a: _ptr(a, ctypes.c_int32)
        R = self.rco.shape[0]
        P, B = self.nrb.shape
        rot_sph = numpy.zeros((R, 4), dtype=self.dtype)
        getattr(L, "oracle_block_spheres" + suf)(
            _ptr(self.x, ct), R, I(self.rco), I(self.btr), I(S.arrays["bt_n_atoms"]),
            _ptr(rot_sph, ct))
        blk_sph = numpy.zeros((P, B, 4), dtype=self.dtype)
        getattr(L, "oracle_rot_block_spheres" + suf)(P, B, _ptr(rot_sph, ct), I(self.nrb),
                                                     I(self.rob), _ptr(blk_sph, ct))
        nbr = numpy.zeros((P, B, B), dtype=numpy.int32)
        getattr(L, "oracle_block_neighbors" + suf)(P, B, I(self.frt), _ptr(blk_sph, ct),
                                                   ct(self.dtype(S.max_dis)), I(nbr))
        n = int(L.oracle_rot_dispatch(P, B, I(nbr), I(self.nrb), I(self.rob), None, 0))
        idx = numpy.full((3, max(n, 1)), -1, dtype=numpy.int32)
        L.oracle_rot_dispatch(P, B, I(nbr), I(self.nrb), I(self.rob), I(idx), n)
        return numpy.ascontiguousarray(idx[:, :n]), rot_sph, blk_sph, nbr

    def score(self, idx, dTdV=None):
        """Energies [3, n] of the listed pairs idx [3, n], or with dTdV [3, n]
        the weighted gradient [NA, 3]."""
        L, S, ct, suf = lib(), self.S, self.ct, self.suf
        I = lambda a: _ptr(a, ctypes.c_int32)
        idx = numpy.ascontiguousarray(numpy.asarray(idx, dtype=numpy.int32))
        n = idx.shape[1]
        NA = self.x.shape[0]
        out = numpy.zeros((3, n), dtype=self.dtype)
        grad = w = None
        if dTdV is not None:
            w = numpy.ascontiguousarray(numpy.asarray(dTdV, dtype=self.dtype).reshape(3, n))
            grad = numpy.zeros((NA, 3), dtype=self.dtype)
        getattr(L, "oracle_rotamer_scores" + suf)(
            ctypes.byref(S.struct), _ptr(self.x, ct), NA, _ptr(self.tp, ct), _ptr(self.gp, ct), n,
            I(idx) if n else None, I(self.rco), I(self.bir), I(self.btr),
            _ptr(w, ct) if w is not None else None,
            _ptr(out, ct) if w is None else None,
            _ptr(grad, ct) if grad is not None else None)
        return grad if dTdV is not None else out


def oracle_rotamer_scores(*args, dTdV=None, dtype=None):
    """Run the oracle on the 27 arguments of ``ljlk_rotamer_scores``
    (compiled.ops.cpp:282-530, ljlk_pose_score.impl.hh:1479-2126).

    Returns (scores [3, n], dispatch_indices [3, n] int32, rot_spheres [R, 4],
    block_spheres [P, B, 4], block_neighbors [P, B, B]); with dTdV [3, n]
    also the weighted gradient [NA, 3] as a sixth item."""
    o = _RotOracle(args, dtype)
    idx, rot_sph, blk_sph, nbr = o.dispatch()
    out = o.score(idx)
    res = (out, idx, rot_sph, blk_sph, nbr)
    return res + (o.score(idx, dTdV),) if dTdV is not None else res


def oracle_rotamer_dispatch(*args, dtype=None):
    """(oracle, indices, rot_spheres, block_spheres, block_neighbors) without
    scoring; ``oracle.score(idx_subset)`` scores any listed pairs."""
    o = _RotOracle(args, dtype)
    return (o,) + o.dispatch()


def oracle_inter_block_bondsep(pbt, block_type, irc):
    """(inter_block_bondsep [P, B, B, C, C], min_block_bondsep [P, B, B]) of
    the connection graph, restating the reference's pose-stack construction
    (pose_stack_builder.py:459-599, 1230-1349)."""
    L = lib()
    bt = _np_i32(block_type)
    ir = _np_i32(irc)
    P, B = bt.shape
    n_conn = _np_i32(pbt.n_conn)
    conn = _np_i32(pbt.conn_atom)
    pd = _np_i32(pbt.ljlk_bond_separation)
    C = conn.shape[1]
    Amax = pd.shape[1]
    ibs = numpy.zeros((P, B, B, C, C), dtype=numpy.int32)
    mins = numpy.zeros((P, B, B), dtype=numpy.int32)
    I = lambda a: _ptr(a, ctypes.c_int32)
    L.oracle_inter_block_bondsep(P, B, C, Amax, I(bt), I(ir), I(n_conn), I(conn), I(pd), I(ibs),
                                 I(mins))
    return ibs, mins


def load_reference_apsp_op():
    """The reference's own all-pairs-shortest-paths op
    (tmol_apsp::apsp_op, tmol/pose/compiled/apsp.cpu.cpp) built by
    build_ref.sh."""
    path = os.path.join(HERE, "_ref", "tmol_apsp_ref.so")
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    if not hasattr(torch.ops, "tmol_apsp") or not hasattr(torch.ops.tmol_apsp, "apsp_op"):
        torch.ops.load_library(path)
    return torch.ops.tmol_apsp.apsp_op


def oracle_pair_separations(args, items):
    """Count-pair separations for every atom pair of each (p, b1, b2) item.
    Returns a list of int arrays [n1, n2] (-1 where not real)."""
    S = _Stack(args)
    L = lib()
    out = []
    bt_n = S.arrays["bt_n_atoms"]
    bt = S.arrays["block_type"]
    for p, b1, b2 in items:
        t1, t2 = bt[p * S.B + b1], bt[p * S.B + b2]
        n1, n2 = bt_n[t1], bt_n[t2]
        m = numpy.zeros((n1, n2), dtype=numpy.int32)
        L.oracle_block_pair_seps(ctypes.byref(S.struct), p, b1, b2, _ptr(m, ctypes.c_int32))
        out.append(m)
    return out


def lj_pair(dist, sep, params_i, params_j, global_params, with_deriv=True,
            dtype=numpy.float64):
    """Single LJ pair: returns (atr, rep, datr/dd, drep/dd)."""
    L = lib()
    suf, ct = ("_d", ctypes.c_double) if dtype == numpy.float64 else ("_f", ctypes.c_float)
    pi = numpy.ascontiguousarray(numpy.asarray(params_i, dtype=dtype))
    pj = numpy.ascontiguousarray(numpy.asarray(params_j, dtype=dtype))
    gp = numpy.ascontiguousarray(numpy.asarray(global_params, dtype=dtype).reshape(-1))
    o = numpy.zeros(4, dtype=dtype)
    getattr(L, "oracle_lj_pair" + suf)(
        ct(dist), int(sep), _ptr(pi, ct), _ptr(pj, ct), _ptr(gp, ct), int(with_deriv),
        _ptr(o, ct),
    )
    return o


def lk_pair(dist, sep, params_i, params_j, global_params, with_deriv=True,
            dtype=numpy.float64):
    """Single LK pair (both directions): returns (V, dV/dd)."""
    L = lib()
    suf, ct = ("_d", ctypes.c_double) if dtype == numpy.float64 else ("_f", ctypes.c_float)
    pi = numpy.ascontiguousarray(numpy.asarray(params_i, dtype=dtype))
    pj = numpy.ascontiguousarray(numpy.asarray(params_j, dtype=dtype))
    gp = numpy.ascontiguousarray(numpy.asarray(global_params, dtype=dtype).reshape(-1))
    o = numpy.zeros(2, dtype=dtype)
    getattr(L, "oracle_lk_pair" + suf)(
        ct(dist), int(sep), _ptr(pi, ct), _ptr(pj, ct), _ptr(gp, ct), int(with_deriv),
        _ptr(o, ct),
    )
    return o


def elec_pair(dist, q1, q2, sep, global_params, with_deriv=True, dtype=numpy.float64):
    """Single elec pair: returns (V, dV/dd); global_params = D, D0, S,
    min_dis, max_dis."""
    L = lib()
    suf, ct = ("_d", ctypes.c_double) if dtype == numpy.float64 else ("_f", ctypes.c_float)
    gp = numpy.ascontiguousarray(numpy.asarray(global_params, dtype=dtype).reshape(-1))
    o = numpy.zeros(2, dtype=dtype)
    getattr(L, "oracle_elec_pair" + suf)(
        ct(dist), ct(q1), ct(q2), int(sep), _ptr(gp, ct), int(with_deriv), _ptr(o, ct))
    return o


class _ElecStack:
    """Host arrays of the elec op's pose-stack and block-type arguments
    (positions 13-22 of ``elec_pose_scores`` / ``elec_rotamer_scores``) as an
    oracle stack: bt_path_distance holds the inter-block representative path
    distances."""

    def __init__(self, args, dtype):
        (coords, rot_coord_offset, pose_ind_for_atom, first_rot_for_block,
         first_rot_block_type, block_ind_for_rot, pose_ind_for_rot,
         block_type_ind_for_rot, n_rots_for_pose, rot_offset_for_pose,
         n_rots_for_block, rot_offset_for_block, max_n_rots_per_pose,
         min_bondsep, inter_block_bondsep, bt_n_atoms, bt_charge, bt_n_conn,
         bt_conn_atoms, bt_inter_rpd, bt_intra_rpd, bt_is_ligand, global_params,
         max_dis, block_pair) = args
        P, B = first_rot_for_block.shape
        T, Am = bt_charge.shape
        if dtype is None:
            dtype = numpy.float64 if coords.dtype == torch.float64 else numpy.float32
        self.dtype = dtype
        self.suf, self.ct = ("_d", ctypes.c_double) if dtype == numpy.float64 else ("_f", ctypes.c_float)
        self.arrays = arrays = {
            "rot_coord_offset": _np_i32(rot_coord_offset),
            "block_type": _np_i32(block_type_ind_for_rot),
            "min_bondsep": _np_i32(min_bondsep),
            "inter_block_bondsep": _np_i32(inter_block_bondsep),
            "bt_n_atoms": _np_i32(bt_n_atoms),
            "bt_n_conn": _np_i32(bt_n_conn),
            "bt_conn_atoms": _np_i32(bt_conn_atoms),
            "bt_path_distance": _np_i32(bt_inter_rpd),
            "bt_is_ligand": _np_i32(bt_is_ligand),
        }
        st = OracleStack()
        st.n_poses, st.max_n_blocks = P, B
        st.n_block_types, st.max_n_block_atoms = T, Am
        st.max_n_tiles = (Am - 1) // 32 + 1
        st.max_n_conn = bt_conn_atoms.shape[1]
        for k, v in arrays.items():
            setattr(st, k, _ptr(v, ctypes.c_int32))
        self.struct = st
        self.P, self.B = P, B
        self.intra = _np_i32(bt_intra_rpd)
        self.x = numpy.ascontiguousarray(coords.detach().cpu().numpy().astype(dtype))
        self.q = numpy.ascontiguousarray(bt_charge.detach().cpu().numpy().astype(dtype))
        self.gp = numpy.ascontiguousarray(
            global_params.detach().cpu().numpy().astype(dtype).reshape(-1))
        self.max_dis = float(max_dis)
        self.block_pair = bool(block_pair)
        self.rest = (first_rot_block_type, block_ind_for_rot, n_rots_for_block,
                     rot_offset_for_block)


def oracle_elec_scores(*args, require_grad=False, dtype=None):
    """Run the elec oracle on the 25 arguments of ``elec_pose_scores``
    (tmol/score/elec/potentials/compiled.ops.cpp:526-586).

    Returns (scores, block_neighbors, dV_dx, spheres): scores is (1, P) or
    (1, P, B, B), block_neighbors (P, B, B) int32, dV_dx (1, NA, 3) or None."""
    E = _ElecStack(args, dtype)
    L, ct, suf, P, B, arrays = lib(), E.ct, E.suf, E.P, E.B, E.arrays
    x = E.x
    NA = x.shape[0]
    spheres = numpy.zeros((P * B, 4), dtype=E.dtype)
    getattr(L, "oracle_block_spheres" + suf)(
        _ptr(x, ct), P * B, arrays["rot_coord_offset"].ctypes.data_as(_P),
        arrays["block_type"].ctypes.data_as(_P), arrays["bt_n_atoms"].ctypes.data_as(_P),
        _ptr(spheres, ct))
    nbr = numpy.zeros((P, B, B), dtype=numpy.int32)
    getattr(L, "oracle_block_neighbors" + suf)(
        P, B, arrays["block_type"].ctypes.data_as(_P), _ptr(spheres, ct),
        ct(E.dtype(E.max_dis)), nbr.ctypes.data_as(_P))
    out = numpy.zeros((1, P, B, B) if E.block_pair else (1, P), dtype=E.dtype)
    dvdx = numpy.zeros((1, NA, 3), dtype=E.dtype) if require_grad else None
    getattr(L, "oracle_elec_scores" + suf)(
        ctypes.byref(E.struct), _ptr(x, ct), NA, _ptr(E.q, ct), E.intra.ctypes.data_as(_P),
        _ptr(E.gp, ct), nbr.ctypes.data_as(_P), int(E.block_pair), int(require_grad),
        _ptr(out, ct), _ptr(dvdx, ct) if dvdx is not None else None)
    return out, nbr, dvdx, spheres


def oracle_elec_rotamer_scores(*args, dTdV=None, dtype=None, indices=None):
    """Run the elec oracle on the 25 arguments of ``elec_rotamer_scores``
    (compiled.ops.cpp:588-650; ElecRotamerScoreDispatch,
    elec_pose_score.impl.hh:1132-1700), the rotamer set in place of the pose
    stack's rotamer tensors.  The dispatch list is the rotamer sphere /
    block-neighbor / dispatch chain of the LJ+LK rotamer oracle with elec's
    max_dis.

    Returns (scores [1, n], dispatch_indices [3, n] int32, block_neighbors
    [P, B, B]); with dTdV [1, n] also the weighted gradient [NA, 3] as a
    fourth item.  ``indices`` [3, n] scores that list instead of the
    dispatch (block_neighbors is then None)."""
    E = _ElecStack(args, dtype)
    L, ct, suf, arrays = lib(), E.ct, E.suf, E.arrays
    frt, bir, nrb, rob = (_np_i32(t) for t in E.rest)
    I = lambda a: _ptr(a, ctypes.c_int32)
    rco, btr = arrays["rot_coord_offset"], arrays["block_type"]
    if indices is not None:
        idx = numpy.ascontiguousarray(numpy.asarray(indices, dtype=numpy.int32))
        n, nbr = idx.shape[1], None
    else:
        idx, nbr = _elec_rot_dispatch(E, L, frt, nrb, rob)
        n = idx.shape[1]
    NA = E.x.shape[0]

    def run(w):
        out = numpy.zeros((1, n), dtype=E.dtype)
        grad = numpy.zeros((NA, 3), dtype=E.dtype) if w is not None else None
        getattr(L, "oracle_elec_rotamer_scores" + suf)(
            ctypes.byref(E.struct), _ptr(E.x, ct), NA, _ptr(E.q, ct), I(E.intra),
            _ptr(E.gp, ct), n, I(idx) if n else None, I(rco), I(bir), I(btr),
            _ptr(w, ct) if w is not None else None,
            _ptr(out, ct) if w is None else None,
            _ptr(grad, ct) if grad is not None else None)
        return grad if w is not None else out

    res = (run(None), idx, nbr)
    if dTdV is not None:
        w = numpy.ascontiguousarray(numpy.asarray(dTdV, dtype=E.dtype).reshape(n))
        res = res + (run(w),)
    return res


def oracle_elec_rotamer_dispatch(*args, dtype=None):
    """(dispatch_indices [3, n] int32, block_neighbors [P, B, B]) of the 25
    arguments of ``elec_rotamer_scores``, without scoring."""
    E = _ElecStack(args, dtype)
    frt, _, nrb, rob = (_np_i32(t) for t in E.rest)
    return _elec_rot_dispatch(E, lib(), frt, nrb, rob)


def _elec_rot_dispatch(E, L, frt, nrb, rob):
    """Rotamer spheres, block spheres, block neighbors at elec's max_dis and
    the dispatch list (the LJ+LK rotamer oracle's chain)."""
    ct, suf, P, B, arrays = E.ct, E.suf, E.P, E.B, E.arrays
    I = lambda a: _ptr(a, ctypes.c_int32)
    rco, btr = arrays["rot_coord_offset"], arrays["block_type"]
    R = rco.shape[0]
    rot_sph = numpy.zeros((R, 4), dtype=E.dtype)
    getattr(L, "oracle_block_spheres" + suf)(
        _ptr(E.x, ct), R, I(rco), I(btr), I(arrays["bt_n_atoms"]), _ptr(rot_sph, ct))
    blk_sph = numpy.zeros((P, B, 4), dtype=E.dtype)
    getattr(L, "oracle_rot_block_spheres" + suf)(P, B, _ptr(rot_sph, ct), I(nrb), I(rob),
                                                 _ptr(blk_sph, ct))
    nbr = numpy.zeros((P, B, B), dtype=numpy.int32)
    getattr(L, "oracle_block_neighbors" + suf)(P, B, I(frt), _ptr(blk_sph, ct),
                                               ct(E.dtype(E.max_dis)), I(nbr))
    n = int(L.oracle_rot_dispatch(P, B, I(nbr), I(nrb), I(rob), None, 0))
    idx = numpy.full((3, max(n, 1)), -1, dtype=numpy.int32)
    L.oracle_rot_dispatch(P, B, I(nbr), I(nrb), I(rob), I(idx), n)
    return numpy.ascontiguousarray(idx[:, :n]), nbr


def reference_op_available():
    """True when oracle/_ref holds the reference's own LJ+LK op.  It runs in a
    child process (oracle/refop.py): tmol_amd registers ops of the same names."""
    return os.path.exists(REF_LIB_PATH)


def load_reference_elec_op(name="elec_pose_scores"):
    """The reference's own CPU elec op built by build_ref.sh
    (``torch.ops.tmol_elec.elec_pose_scores``, or ``elec_rotamer_scores``)."""
    path = os.path.join(HERE, "_ref", "tmol_elec_ref.so")
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    if not hasattr(torch.ops, "tmol_elec") or not hasattr(torch.ops.tmol_elec, "elec_pose_scores"):
        torch.ops.load_library(path)
    return getattr(torch.ops.tmol_elec, name)


def pose_oracle_evaluator(poses, weights=(1.0, 0.55, 1.0), threads=8, dtype=numpy.float64):
    """Energy + weighted gradient of a list of poses by the oracle, one pose
    per call (poses are independent), spread over ``threads`` host threads.

    The parameter vector is the poses' atoms concatenated in order (the
    ragged stack's layout), flattened; returns (E [P] float64, dE/dx [3 NA])
    as torch tensors.  Used to drive the minimizer from the oracle."""
    from concurrent.futures import ThreadPoolExecutor

    from tmol_amd.pose import PoseStack
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

    tdt = torch.float64 if dtype == numpy.float64 else torch.float32
    term = LJLKEnergyTerm()
    per_pose = []
    offset = 0
    for pose in poses:
        a = pose_op_args(PoseStack([pose], dtype=tdt), term)
        a[23] = a[23].to(tdt)
        a[24] = a[24].to(tdt)
        per_pose.append((offset, pose.n_atoms, a))
        offset += pose.n_atoms
    w = numpy.asarray(weights, dtype=numpy.float64)
    pool = ThreadPoolExecutor(max_workers=threads)

    prepared = [(off, n, _Stack(a)) for off, n, a in per_pose]

    def one(item, xn):
        off, n, S = item
        c = numpy.ascontiguousarray(xn[3 * off: 3 * (off + n)].reshape(n, 3).astype(dtype))
        s, _, dv, _ = _scores(S, c, True, dtype)
        e = float((w * s[:, 0].astype(numpy.float64)).sum())
        g = (w[:, None, None] * dv.astype(numpy.float64)).sum(0).reshape(-1)
        return e, g

    def evaluate(x):
        xn = x.detach().cpu().numpy().astype(numpy.float64)
        res = list(pool.map(lambda it: one(it, xn), prepared))
        E = torch.tensor([r[0] for r in res], dtype=torch.float64)
        G = torch.from_numpy(numpy.concatenate([r[1] for r in res]))
        return E.to(x.device), G.to(device=x.device, dtype=x.dtype)

    return evaluate


# ---------------------------------------------------------------- lk_ball

class OracleLKBStack(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "n_poses", "max_n_blocks", "n_block_types", "max_n_block_atoms", "max_n_tiles",
        "max_n_conn", "max_n_all_bonds")] + [(n, _P) for n in (
            "rot_coord_offset", "block_type", "irc", "min_bondsep", "inter_block_bondsep",
            "bt_n_atoms", "bt_n_conn", "bt_conn_atoms", "bt_all_bonds", "bt_atom_all_bond_ranges",
            "tile_n_donH", "tile_n_acc", "tile_donH_inds", "tile_donH_hvy_inds", "tile_which_donH",
            "tile_acc_inds", "tile_acc_hyb", "tile_acc_nH", "is_hydrogen", "tile_n_polar",
            "tile_n_occ", "tile_pol_occ_inds", "path_distance")] + [
        ("n_tors", ctypes.c_int32 * 3)]


class _LKBStack:
    """The lk_ball oracle's view of the two reference ops' arguments: the 33
    of gen_pose_waters (lk_ball_energy_term.py:244-268) and, for scoring, the
    28 of lk_ball_pose_score (:272-296)."""

    def __init__(self, wargs, sargs=None):
        (coords, rco, _pia, frfb, frbt, _bifr, _pifr, btifr, _nrp, _rop, _nrb, _rob, _mx,
         irc, bt_n_atoms, bt_n_conn, bt_conn_atoms, _nab, all_bonds, ranges, t_ndonH, t_nacc,
         t_donH, t_donHhvy, t_which, t_acc, t_hyb, t_nH, is_h, wg, sp2, sp3, ring) = wargs
        P, B = frfb.shape
        a = {
            "rot_coord_offset": _np_i32(rco), "block_type": _np_i32(btifr), "irc": _np_i32(irc),
            "bt_n_atoms": _np_i32(bt_n_atoms), "bt_n_conn": _np_i32(bt_n_conn),
            "bt_conn_atoms": _np_i32(bt_conn_atoms), "bt_all_bonds": _np_i32(all_bonds),
            "bt_atom_all_bond_ranges": _np_i32(ranges), "tile_n_donH": _np_i32(t_ndonH),
            "tile_n_acc": _np_i32(t_nacc), "tile_donH_inds": _np_i32(t_donH),
            "tile_donH_hvy_inds": _np_i32(t_donHhvy), "tile_which_donH": _np_i32(t_which),
            "tile_acc_inds": _np_i32(t_acc), "tile_acc_hyb": _np_i32(t_hyb),
            "tile_acc_nH": _np_i32(t_nH), "is_hydrogen": _np_i32(is_h),
        }
        self.block_pair = False
        if sargs is not None:
            (_c, _rco, _pia, _frfb, _frbt, _bifr, _pifr, _btifr, _nrp, _rop, _nrb, _rob, _mx,
             _irc, mbs, ibs, _bna, _bnc, _bca, t_npol, t_nocc, t_poi, t_par, pd, gp, max_dis,
             _w, block_pair) = sargs
            a.update({"min_bondsep": _np_i32(mbs), "inter_block_bondsep": _np_i32(ibs),
                      "tile_n_polar": _np_i32(t_npol), "tile_n_occ": _np_i32(t_nocc),
                      "tile_pol_occ_inds": _np_i32(t_poi), "path_distance": _np_i32(pd)})
            self.tile_params, self.gp = t_par, gp
            self.max_dis, self.block_pair = float(max_dis), bool(block_pair)
        self.arrays = a
        st = OracleLKBStack()
        st.n_poses, st.max_n_blocks = P, B
        st.n_block_types = bt_n_atoms.shape[0]
        st.max_n_block_atoms = is_h.shape[1]
        st.max_n_tiles = t_ndonH.shape[1]
        st.max_n_conn = bt_conn_atoms.shape[1]
        st.max_n_all_bonds = all_bonds.shape[1]
        for k, v in a.items():
            setattr(st, k, _ptr(v, ctypes.c_int32))
        st.n_tors[0], st.n_tors[1], st.n_tors[2] = len(sp2), len(sp3), len(ring)
        self.struct = st
        self.P, self.B = P, B
        self.coords, self.wg, self.tors = coords, wg, (sp2, sp3, ring)
        self.bt_n_atoms = a["bt_n_atoms"]

    def host(self, t, dtype):
        return numpy.ascontiguousarray(t.detach().cpu().numpy().astype(dtype))


def _lkb_ct(dtype):
    return ("_d", ctypes.c_double) if dtype == numpy.float64 else ("_f", ctypes.c_float)


def oracle_lkb_waters(*wargs, dtype=None):
    """gen_pose_waters on the oracle: [NA, 4, 3], NaN where no water."""
    S = _LKBStack(wargs)
    if dtype is None:
        dtype = numpy.float64 if S.coords.dtype == torch.float64 else numpy.float32
    suf, ct = _lkb_ct(dtype)
    X = S.host(S.coords, dtype)
    wg = S.host(S.wg, dtype).reshape(-1)
    tors = [S.host(t, dtype) for t in S.tors]
    W = numpy.zeros((X.shape[0], 4, 3), dtype=dtype)
    getattr(lib(), "oracle_lkb_waters" + suf)(
        ctypes.byref(S.struct), _ptr(X, ct), X.shape[0], _ptr(wg, ct),
        *[_ptr(t, ct) for t in tors], _ptr(W, ct))
    return W


def oracle_lkb_waters_backward(wargs, dW, dtype=None):
    """dE/dx of the water builders for dE/dW [NA, 4, 3]."""
    S = _LKBStack(wargs)
    if dtype is None:
        dtype = numpy.float64 if S.coords.dtype == torch.float64 else numpy.float32
    suf, ct = _lkb_ct(dtype)
    X = S.host(S.coords, dtype)
    wg = S.host(S.wg, dtype).reshape(-1)
    tors = [S.host(t, dtype) for t in S.tors]
    g = numpy.ascontiguousarray(numpy.asarray(dW, dtype=dtype))
    dX = numpy.zeros((X.shape[0], 3), dtype=dtype)
    getattr(lib(), "oracle_lkb_waters_backward" + suf)(
        ctypes.byref(S.struct), _ptr(X, ct), _ptr(wg, ct), *[_ptr(t, ct) for t in tors],
        _ptr(g, ct), _ptr(dX, ct))
    return dX


def oracle_lkb_score(wargs, sargs, waters=None, dTdV=None, dtype=None):
    """lk_ball_pose_score on the oracle.  Returns (scores [4, P] or
    [4, P, B, B], block_neighbors, dX [NA, 3] or None, dW [NA, 4, 3] or
    None); with dTdV (shape of scores) the derivatives are the weighted
    backward of the op, dX without the water chain rule."""
    S = _LKBStack(wargs, sargs)
    if dtype is None:
        dtype = numpy.float64 if S.coords.dtype == torch.float64 else numpy.float32
    suf, ct = _lkb_ct(dtype)
    L = lib()
    X = S.host(S.coords, dtype)
    NA = X.shape[0]
    if waters is None:
        waters = S.host(sargs[26], dtype)
    W = numpy.ascontiguousarray(numpy.asarray(waters, dtype=dtype).reshape(NA, 4, 3))
    PB = S.P * S.B
    spheres = numpy.zeros((PB, 4), dtype=dtype)
    getattr(L, "oracle_block_spheres" + suf)(
        _ptr(X, ct), PB, _ptr(S.arrays["rot_coord_offset"], ctypes.c_int32),
        _ptr(S.arrays["block_type"], ctypes.c_int32), _ptr(S.bt_n_atoms, ctypes.c_int32),
        _ptr(spheres, ct))
    nbr = numpy.zeros((S.P, S.B, S.B), dtype=numpy.int32)
    getattr(L, "oracle_block_neighbors" + suf)(
        S.P, S.B, _ptr(S.arrays["block_type"], ctypes.c_int32), _ptr(spheres, ct),
        ct(dtype(S.max_dis)), _ptr(nbr, ctypes.c_int32))
    shape = (4, S.P, S.B, S.B) if S.block_pair else (4, S.P)
    out = numpy.zeros(shape, dtype=dtype)
    tp = S.host(S.tile_params, dtype)
    gp = S.host(S.gp, dtype).reshape(-1)
    dX = dW = wts = None
    if dTdV is not None:
        wts = numpy.ascontiguousarray(numpy.asarray(dTdV, dtype=dtype).reshape(shape))
        dX = numpy.zeros((NA, 3), dtype=dtype)
        dW = numpy.zeros((NA, 4, 3), dtype=dtype)
    getattr(L, "oracle_lkb_score" + suf)(
        ctypes.byref(S.struct), _ptr(X, ct), _ptr(W, ct), NA, _ptr(tp, ct), _ptr(gp, ct),
        _ptr(nbr, ctypes.c_int32), int(S.block_pair), _ptr(out, ct),
        _ptr(wts, ct) if wts is not None else None, _ptr(dX, ct) if dX is not None else None,
        _ptr(dW, ct) if dW is not None else None)
    return out, nbr, dX, dW


def oracle_lkb_pair(I, J, WI, WJ, sep, params_i, params_j, global_params, deriv=False,
                    dtype=numpy.float64):
    """lk_ball_score::V (and dV) of one polar / occluder pair: E[4] and, with
    deriv, (dI [4, 3], dJ [4, 3], dWI [4, 4, 3], dWJ [4, 4, 3])."""
    suf, ct = _lkb_ct(dtype)
    f = lambda x, n: numpy.ascontiguousarray(numpy.asarray(x, dtype=dtype).reshape(n))  # noqa
    I, J = f(I, 3), f(J, 3)
    WI, WJ = f(WI, 12), f(WJ, 12)
    pi, pj, gp = f(params_i, 9), f(params_j, 9), f(global_params, 5)
    E = numpy.zeros(4, dtype=dtype)
    d = numpy.zeros(4 * 3 * 2 + 4 * 12 * 2, dtype=dtype) if deriv else None
    getattr(lib(), "oracle_lkb_pair" + suf)(
        _ptr(I, ct), _ptr(J, ct), _ptr(WI, ct), _ptr(WJ, ct), int(sep), _ptr(pi, ct),
        _ptr(pj, ct), _ptr(gp, ct), _ptr(E, ct), _ptr(d, ct) if d is not None else None)
    if not deriv:
        return E
    return E, (d[:12].reshape(4, 3), d[12:24].reshape(4, 3), d[24:72].reshape(4, 4, 3),
               d[72:].reshape(4, 4, 3))


def load_reference_lk_ball_ops():
    """The reference's own CPU lk_ball ops built by build_ref.sh
    (``torch.ops.tmol_lk_ball``: gen_pose_waters, lk_ball_pose_score)."""
    path = os.path.join(HERE, "_ref", "tmol_lk_ball_ref.so")
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    if not hasattr(torch.ops, "tmol_lk_ball") or not hasattr(torch.ops.tmol_lk_ball,
                                                              "gen_pose_waters"):
        torch.ops.load_library(path)
    return torch.ops.tmol_lk_ball


def oracle_lkb_build_water(kind, P, G, GG=None, dist=2.65, angle=0.0, tors=0.0, g=None,
                           dtype=numpy.float64):
    """One water: kind 0 donor (P = D, G = H), otherwise build_coordinate
    (P = A, G = B, GG = B0).  With g = dE/dW also returns (dE/dP, dE/dG,
    dE/dGG)."""
    suf, ct = _lkb_ct(dtype)
    f = lambda x: numpy.ascontiguousarray(numpy.asarray(x, dtype=dtype).reshape(3))  # noqa
    P, G = f(P), f(G)
    GG = f(GG if GG is not None else numpy.zeros(3))
    W = numpy.zeros(3, dtype=dtype)
    outs = [numpy.zeros(3, dtype=dtype) for _ in range(3)]
    gg = f(g) if g is not None else None
    getattr(lib(), "oracle_lkb_build_water" + suf)(
        int(kind), _ptr(P, ct), _ptr(G, ct), _ptr(GG, ct), ct(dist), ct(angle), ct(tors),
        _ptr(W, ct), _ptr(gg, ct) if gg is not None else None, *[_ptr(o, ct) for o in outs])
    return W if g is None else (W, outs)


def _waters4(W, dtype):
    """[n <= 4, 3] waters -> [4, 3], NaN rows appended"""
    W = numpy.asarray(W, dtype=dtype).reshape(-1, 3)
    out = numpy.full((4, 3), numpy.nan, dtype=dtype)
    out[: W.shape[0]] = W
    return out


def oracle_lkb_fraction(WI, J, lj_radius_j, deriv=False, dtype=numpy.float64):
    """lk_fraction::V (and dV: (dWI [4, 3], dJ [3]))."""
    suf, ct = _lkb_ct(dtype)
    WI = _waters4(WI, dtype)
    J = numpy.ascontiguousarray(numpy.asarray(J, dtype=dtype).reshape(3))
    dWI, dJ = numpy.zeros((4, 3), dtype=dtype), numpy.zeros(3, dtype=dtype)
    v = getattr(lib(), "oracle_lkb_fraction" + suf)(
        _ptr(WI, ct), _ptr(J, ct), ct(lj_radius_j), _ptr(dWI, ct) if deriv else None,
        _ptr(dJ, ct) if deriv else None)
    return (v, dWI, dJ) if deriv else v


def oracle_lkb_bridge(I, J, WI, WJ, water_dist, deriv=False, dtype=numpy.float64):
    """lk_bridge_fraction::V (and dV: (dI, dJ, dWI, dWJ))."""
    suf, ct = _lkb_ct(dtype)
    f = lambda x: numpy.ascontiguousarray(numpy.asarray(x, dtype=dtype).reshape(3))  # noqa
    I, J = f(I), f(J)
    WI, WJ = _waters4(WI, dtype), _waters4(WJ, dtype)
    outs = [numpy.zeros(3, dtype=dtype), numpy.zeros(3, dtype=dtype),
            numpy.zeros((4, 3), dtype=dtype), numpy.zeros((4, 3), dtype=dtype)]
    v = getattr(lib(), "oracle_lkb_bridge" + suf)(
        _ptr(I, ct), _ptr(J, ct), _ptr(WI, ct), _ptr(WJ, ct), ct(water_dist),
        *[_ptr(o, ct) if deriv else None for o in outs])
    return (v, *outs) if deriv else v


# ---------------------------------------------------------------- hbond

class OracleHBStack(ctypes.Structure):
    _fields_ = [("base", OracleLKBStack), ("tile_don_type", _P), ("tile_acc_type", _P),
                ("n_don_types", ctypes.c_int32), ("n_acc_types", ctypes.c_int32)]


class _HBStack:
    """The hbond oracle's view of the reference ops' arguments: the 20 of
    gen_hbond_bases (hbond_energy_term.py:76-97) and, for scoring, the 37 of
    hbond_pose_scores (:98-132)."""

    def __init__(self, bargs, sargs=None):
        (coords, rco, frfb, _frbt, _bifr, _pifr, btifr, irc, bt_n_atoms, bt_n_conn,
         bt_conn_atoms, _nab, all_bonds, ranges, t_ndonH, t_nacc, t_donH, t_acc, t_hyb,
         is_h) = bargs
        P, B = frfb.shape
        zero_t = numpy.zeros(1, dtype=numpy.int32)
        a = {
            "rot_coord_offset": _np_i32(rco), "block_type": _np_i32(btifr), "irc": _np_i32(irc),
            "bt_n_atoms": _np_i32(bt_n_atoms), "bt_n_conn": _np_i32(bt_n_conn),
            "bt_conn_atoms": _np_i32(bt_conn_atoms), "bt_all_bonds": _np_i32(all_bonds),
            "bt_atom_all_bond_ranges": _np_i32(ranges), "tile_n_donH": _np_i32(t_ndonH),
            "tile_n_acc": _np_i32(t_nacc), "tile_donH_inds": _np_i32(t_donH),
            "tile_acc_inds": _np_i32(t_acc), "tile_acc_hyb": _np_i32(t_hyb),
            "is_hydrogen": _np_i32(is_h),
        }
        hb = {}
        self.block_pair = False
        if sargs is not None:
            (_c, _rco, _pia, _frfb, _frbt, _bifr, _pifr, _btifr, _nrp, _rop, _nrb, _rob, _mx,
             _irc, mbs, ibs, _bna, _bnc, _bca, _nab2, _ab2, _rng2, pd, _tnd, _tna, _tdi, _tai,
             t_dtype, t_atype, _thyb, _ish, pair_params, pair_polys, gp, derived, dinds,
             block_pair) = sargs
            a.update({"min_bondsep": _np_i32(mbs), "inter_block_bondsep": _np_i32(ibs),
                      "path_distance": _np_i32(pd)})
            hb = {"tile_don_type": _np_i32(t_dtype), "tile_acc_type": _np_i32(t_atype)}
            self.pair_params, self.pair_polys, self.gp = pair_params, pair_polys, gp
            self.derived, self.dinds = derived, dinds
            self.block_pair = bool(block_pair)
            self.n_types = tuple(pair_params.shape[:2])
        self.arrays = dict(a, **hb)
        st = OracleHBStack()
        b = st.base
        b.n_poses, b.max_n_blocks = P, B
        b.n_block_types = bt_n_atoms.shape[0]
        b.max_n_block_atoms = is_h.shape[1]
        b.max_n_tiles = t_ndonH.shape[1]
        b.max_n_conn = bt_conn_atoms.shape[1]
        b.max_n_all_bonds = all_bonds.shape[1]
        for k, v in a.items():
            setattr(b, k, _ptr(v, ctypes.c_int32))
        for k in ("tile_donH_hvy_inds", "tile_which_donH", "tile_acc_nH", "tile_n_polar",
                  "tile_n_occ", "tile_pol_occ_inds"):
            setattr(b, k, _ptr(zero_t, ctypes.c_int32))
        self._zero = zero_t
        for k, v in hb.items():
            setattr(st, k, _ptr(v, ctypes.c_int32))
        if sargs is not None:
            st.n_don_types, st.n_acc_types = self.n_types
        self.struct = st
        self.P, self.B = P, B
        self.coords = coords
        self.bt_n_atoms = a["bt_n_atoms"]

    def host(self, t, dtype):
        return numpy.ascontiguousarray(t.detach().cpu().numpy().astype(dtype))


def oracle_hb_bases(*bargs, dtype=None):
    """gen_hbond_bases on the oracle: (derived coords [NA, 3, 3], source
    atoms [NA, 3] int32)."""
    S = _HBStack(bargs)
    if dtype is None:
        dtype = numpy.float64 if S.coords.dtype == torch.float64 else numpy.float32
    suf, ct = _lkb_ct(dtype)
    X = S.host(S.coords, dtype)
    NA = X.shape[0]
    D = numpy.zeros((NA, 3, 3), dtype=dtype)
    I = numpy.zeros((NA, 3), dtype=numpy.int32)
    getattr(lib(), "oracle_hb_bases" + suf)(ctypes.byref(S.struct), _ptr(X, ct), NA,
                                            _ptr(D, ct), _ptr(I, ctypes.c_int32))
    return D, I


def oracle_hb_score(sargs, bargs, dTdV=None, dtype=None, derived=None, inds=None):
    """hbond_pose_scores on the oracle.  Returns (scores [1, P] or
    [1, P, B, B], block_neighbors [P, B, B], dX [NA, 3] or None); with dTdV
    (shape of scores) dX is the op's weighted backward."""
    S = _HBStack(bargs, sargs)
    if dtype is None:
        dtype = numpy.float64 if S.coords.dtype == torch.float64 else numpy.float32
    suf, ct = _lkb_ct(dtype)
    L = lib()
    X = S.host(S.coords, dtype)
    NA = X.shape[0]
    Dv = numpy.ascontiguousarray(numpy.asarray(
        derived if derived is not None else S.host(S.derived, dtype), dtype=dtype))
    Iv = numpy.ascontiguousarray(numpy.asarray(
        inds if inds is not None else S.dinds.detach().cpu().numpy(), dtype=numpy.int32))
    PB = S.P * S.B
    spheres = numpy.zeros((PB, 4), dtype=dtype)
    getattr(L, "oracle_block_spheres" + suf)(
        _ptr(X, ct), PB, _ptr(S.arrays["rot_coord_offset"], ctypes.c_int32),
        _ptr(S.arrays["block_type"], ctypes.c_int32), _ptr(S.bt_n_atoms, ctypes.c_int32),
        _ptr(spheres, ct))
    nbr = numpy.zeros((S.P, S.B, S.B), dtype=numpy.int32)
    getattr(L, "oracle_block_neighbors" + suf)(
        S.P, S.B, _ptr(S.arrays["block_type"], ctypes.c_int32), _ptr(spheres, ct),
        ct(dtype(5.5)), _ptr(nbr, ctypes.c_int32))
    shape = (1, S.P, S.B, S.B) if S.block_pair else (1, S.P)
    out = numpy.zeros(shape, dtype=dtype)
    pp = S.host(S.pair_params, dtype)
    polys = S.host(S.pair_polys, numpy.float64)
    gp = S.host(S.gp, dtype).reshape(-1)
    dX = wts = None
    if dTdV is not None:
        wts = numpy.ascontiguousarray(numpy.asarray(dTdV, dtype=dtype).reshape(shape))
        dX = numpy.zeros((NA, 3), dtype=dtype)
    getattr(L, "oracle_hb_score_pose" + suf)(
        ctypes.byref(S.struct), _ptr(X, ct), NA, _ptr(Dv, ct), _ptr(Iv, ctypes.c_int32),
        _ptr(pp, ct), _ptr(polys, ctypes.c_double), _ptr(gp, ct), _ptr(nbr, ctypes.c_int32),
        int(S.block_pair), _ptr(out, ct), _ptr(wts, ct) if wts is not None else None,
        _ptr(dX, ct) if dX is not None else None)
    return out, nbr, dX


def oracle_hb_point(D, H, A, B, B0, pair_params, polys, global_params, deriv=False,
                    dtype=numpy.float64):
    """The five-atom hbond energy of one donor / acceptor geometry; with
    deriv also its gradient [5, 3] (D, H, A, B, B0)."""
    suf, ct = _lkb_ct(dtype)
    f = lambda x, n: numpy.ascontiguousarray(numpy.asarray(x, dtype=dtype).reshape(n))  # noqa
    pts = [f(x, 3) for x in (D, H, A, B, B0)]
    pp = f(pair_params, 3)
    P = numpy.ascontiguousarray(numpy.asarray(polys, dtype=numpy.float64).reshape(48))
    gp = f(global_params, 6)
    g = numpy.zeros(15, dtype=dtype) if deriv else None
    E = getattr(lib(), "oracle_hb_score" + suf)(
        *[_ptr(x, ct) for x in pts], _ptr(pp, ct), _ptr(P, ctypes.c_double), _ptr(gp, ct),
        _ptr(g, ct) if g is not None else None)
    return (E, g.reshape(5, 3)) if deriv else E


def load_reference_hbond_ops():
    """The reference's own CPU hbond ops built by build_ref.sh
    (``torch.ops.tmol_hbond``: gen_hbond_bases, hbond_pose_scores)."""
    path = os.path.join(HERE, "_ref", "tmol_hbond_ref.so")
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    if not hasattr(torch.ops, "tmol_hbond") or not hasattr(torch.ops.tmol_hbond,
                                                            "gen_hbond_bases"):
        torch.ops.load_library(path)
    return torch.ops.tmol_hbond
