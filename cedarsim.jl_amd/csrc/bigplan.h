// bigplan.h — symbolic plan of the supernodal multifrontal path, used for
// circuits too large for the entry-level LU schedule (a power-grid mesh of
// 10^5..10^6 unknowns).
//
// Pivot order: the pivot matching of build_plan (singletons, maximum
// transversal) pairs every column with a row; the matched matrix M has a
// zero-free diagonal and is ordered symmetrically by nested dissection of
// the graph of M + M^T (recursive breadth-first bisection), with hub
// vertices (supply nodes touching thousands of devices) ordered last.
//
// Factor: elimination tree, column structures of L (= structure of U^T,
// the pattern of M + M^T being symmetric), fundamental supernodes merged
// with small descendants into fronts.  A front of supernode s is a dense
// m x m column-major block in one pool: rows/columns = its k pivots, then
// its r update rows (ascending pivot index).  Partial LU of a front leaves
// L11\U11, L21, U12 in place and the Schur complement (contribution block,
// CB) in the trailing r x r block, which the parent gathers.  Fronts are
// processed level by level (height in the assembly tree), all fronts of a
// level in one launch.
#pragma once
#include "plan.h"

namespace cedar {

struct BigPlan {
    Plan base;                          // stamps, natural pattern, devices, sources
    int32_t n = 0;
    // pivot k: natural row rownat[k], natural column colnat[k]
    std::vector<int32_t> rownat, colnat, prow, pcol;   // prow: natural row -> k, pcol: natural col -> k
    int32_t n_hub = 0;
    // supernodes (fronts), in pivot order
    int32_t ns = 0;
    std::vector<int32_t> sn_first;      // [ns+1] pivots [sn_first[s], sn_first[s+1])
    std::vector<int32_t> sn_m;          // front size m = k + r
    std::vector<int32_t> sn_parent;     // parent supernode, -1 for a root
    std::vector<int32_t> sn_rptr, sn_rows;   // update rows (pivot indices), ascending
    std::vector<int64_t> sn_off;        // [ns+1] front offset in the pool (doubles)
    std::vector<int32_t> sn_level;
    std::vector<int32_t> lev_ptr, lev_sn;    // supernodes by level (ascending level)
    std::vector<int32_t> ch_ptr, ch_sn;      // children of each supernode (ascending)
    // per supernode as a child: position of each of its update rows in the
    // parent's front (relmap), and the inverse map over the parent's front
    // rows (child CB index or -1)
    std::vector<int32_t> rel_ptr, rel;
    std::vector<int64_t> inv_ptr;
    std::vector<int32_t> inv;
    // natural CSR position -> pool index; positions grouped by the level of
    // their front
    std::vector<int64_t> nat_dest;
    std::vector<int32_t> olev_ptr, olev_pos;
    std::vector<int32_t> osn_ptr, osn_pos;   // the same positions grouped by front
    int64_t pool = 0;                   // doubles
    int64_t inv_total = 0;
    double flops = 0.0;                 // partial-LU flops of one factorization
    int64_t nnz_factor = 0;             // entries of L+U held in fronts
    int32_t max_m = 0, max_k = 0;
};

// Large-circuit plan.  Returns CS_OK or a CS_E_* code with a message in err.
int build_bigplan(const cs_netlist* nl, BigPlan& B, std::string& err);

}  // namespace cedar
