"""Term scoring modules (tmol/score/common/scoring_module.py:7-199): freeze
the pose-stack (or rotamer-set) index tensors and term attributes once, call
the op with the flattened coordinates on every evaluation."""

import torch


class TermPoseScoringModule(torch.nn.Module):
    def __init__(self, classname, pose_stack, term_parameters, term_score_poses,
                 block_pair_scoring):
        super().__init__()
        self.classname = classname
        self.term_score_poses = term_score_poses
        ps = pose_stack
        common = [
            ps.rot_coord_offset,
            ps.pose_ind_for_atom,
            ps.first_rot_for_block,
            ps.block_type_ind,
            ps.block_ind_for_rot,
            ps.pose_ind_for_rot,
            ps.block_type_ind_for_rot,
            ps.n_rots_for_pose,
            ps.rot_offset_for_pose,
            ps.n_rots_for_block,
            ps.rot_offset_for_block,
            ps.max_n_rots_per_pose,
        ]
        tail = list(common) + list(term_parameters)
        self._tail_f32 = tuple(tail) + (block_pair_scoring,)
        # float64 tail: float tensors promoted (tmol/score/common/convert_float64.py)
        self._tail_f64 = tuple(
            t.to(torch.float64) if torch.is_tensor(t) and t.is_floating_point() else t
            for t in tail
        ) + (block_pair_scoring,)
        self.n_poses = ps.first_rot_for_block.shape[0]
        self.max_n_blocks = ps.first_rot_for_block.shape[1]

    def forward(self, coords):
        flat = coords.flatten(start_dim=0, end_dim=-2)
        tail = self._tail_f64 if coords.dtype == torch.float64 else self._tail_f32
        scores, _ = self.term_score_poses(flat, *tail)
        return scores


class TermWholePoseScoringModule(TermPoseScoringModule):
    def __init__(self, classname, pose_stack, term_parameters, term_score_poses):
        super().__init__(classname, pose_stack, term_parameters, term_score_poses, False)


class TermBlockPairScoringModule(TermPoseScoringModule):
    def __init__(self, classname, pose_stack, term_parameters, term_score_poses):
        super().__init__(classname, pose_stack, term_parameters, term_score_poses, True)


class TermRotamerScoringModule(torch.nn.Module):
    """Rotamer-pair scoring of one term (scoring_module.py:133-199): the
    rotamer set's index tensors in the pose stack's places, block-pair output
    on.  ``forward(coords)`` returns (scores [n_subterms, nnz], indices
    [3, nnz] = pose, rot1, rot2) without building a sparse tensor."""

    def __init__(self, classname, rotamer_set, term_parameters, term_score_poses):
        super().__init__()
        self.classname = classname
        self.term_score_poses = term_score_poses
        rs = rotamer_set

        def i32(x):
            return x if isinstance(x, int) else x.to(torch.int32)

        common = [i32(t) for t in [
            rs.coord_offset_for_rot,
            rs.pose_ind_for_atom,
            rs.rot_offset_for_block,
            rs.first_rot_block_type,
            rs.block_ind_for_rot,
            rs.pose_for_rot,
            rs.block_type_ind_for_rot,
            rs.n_rots_for_pose,
            rs.rot_offset_for_pose,
            rs.n_rots_for_block,
            rs.rot_offset_for_block,
            int(rs.max_n_rots_per_pose),
        ]]
        tail = list(common) + list(term_parameters)
        self._tail_f32 = tuple(tail) + (True,)
        self._tail_f64 = tuple(
            t.to(torch.float64) if torch.is_tensor(t) and t.is_floating_point() else t
            for t in tail
        ) + (True,)
        self.n_poses = rs.n_rots_for_pose.shape[0]
        self.n_rots = rs.coord_offset_for_rot.shape[0]

    def forward(self, coords):
        flat = coords.flatten(start_dim=0, end_dim=-2)
        tail = self._tail_f64 if coords.dtype == torch.float64 else self._tail_f32
        scores, indices = self.term_score_poses(flat, *tail)
        return scores, indices
