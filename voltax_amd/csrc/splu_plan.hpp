// Symbolic analysis of the supernodal (multifrontal) sparse LU, host side.
//
// Same job as the reference's Plan (voltax/sparse.py:257-535): zero-diagonal
// unknowns (voltage-source currents) are glued to a partner terminal node
// so they are always pivoted inside one dense block; a nested-dissection
// ordering of the quotient graph gives supernodes, an elimination tree and
// level sets of independent supernodes.  The numeric layout differs: every
// supernode s owns a dense frontal matrix over (pivots K_s, border R_s);
// children hand their Schur complements to the parent by an extend-add
// through precomputed index maps (multifrontal), so no two supernodes ever
// write the same memory and every sum has a fixed order.
#pragma once
#include <cstdint>
#include <vector>

#include "topology.hpp"

namespace vxm {

struct SNode {
  int k = 0, r = 0;       // pivots, border unknowns
  int parent = -1;        // supernode index (postorder) or -1
  int level = 0;          // 0 = leaf
  int64_t front_off = 0;  // offset of the (k+r)^2 front in the factor store
  int piv_off = 0;        // offset of k pivot entries
  int idx_off = 0;        // offset of the k+r global unknown indices
  int cmap_off = 0;       // offset of r entries: border position in parent's front
  int child_off = 0;      // offset into child list
  int n_child = 0;
  int aent_off = 0;       // offset into the A-entry list
  int n_aent = 0;
  int uvec_off = 0;       // offset of the r-vector handed to the parent in solves
};

struct SPlanHost {
  int S = 0, nnz = 0;
  std::vector<SNode> sn;           // postorder
  std::vector<int> idx;            // front index lists (K then R)
  std::vector<int> cmap;           // child border -> parent front position
  std::vector<int> children;       // children lists
  std::vector<int> aslot, apos;    // A entry: CSR slot -> front position (row*nf+col)
  std::vector<int> level_ptr;      // supernodes of level l: level_sn[level_ptr[l]..)
  std::vector<int> level_sn;       // within a level: fronts largest first
  // level_cls[5l + c] .. level_cls[5l + c + 1]: the level's fronts of size class c
  // (c = 0: k + r > 64, 1: 33..64, 2: 17..32, 3: <= 16), as level_sn offsets
  std::vector<int> level_cls;
  int64_t front_total = 0;         // doubles per instance
  int piv_total = 0;
  int r_total = 0;                 // sum of border sizes
  int max_front = 0, max_k = 0;
  // schedule knob (set by the autotuner): fronts of 33..64 unknowns are
  // factored by the register-resident wavefront kernel (0) or by the blocked
  // 256-thread LDS kernel (1)
  int mid_blocked = 0;
  double lu_flops = 0.0, factor_entries = 0.0;
};

// leaf: largest ND leaf (unknowns) kept as one supernode.
// compact: the grouping of vxm_splu_create_ex's VXM_SPLU_COMPACT (else the
// reference's pivot groups, voltax/sparse.py:257-305).
int build_splu_plan(const Topology& t, int leaf, SPlanHost* out, bool compact = false);

}  // namespace vxm
