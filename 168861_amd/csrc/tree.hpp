// Exact solver for the tree-structured backward-Euler systems (TPS: tree partition
// solve), shared by the single-tree and the batched implicit kernels.
//
// The assembled matrix of a neurite is a forest: each non-identity row i couples
// to at most one "parent" column p(i) < i (its mother, unless folded away at a
// boundary, solverIm-gpu.cpp:216-231) and to its children (right/left daughters).
// The sequential exact method is Hines elimination (children into parents, then
// back-substitution). TPS exposes parallelism as follows:
//   * chains (runs of consecutive rows where each row's only child is the next
//     row) are cut into chunks of at most C rows;
//   * each chunk eliminates its interior rows into its two end rows (a forward
//     sweep, plus a backward one for the reduced rows' coefficients when the chunk's
//     values changed; one worker per chunk) giving a reduced system on the chunk
//     ends with a direct end-to-end coupling; the reduced system is again a forest
//     of chains, so this repeats until every chain has at most two nodes;
//   * the final compressed forest is eliminated by chain tasks ordered by height (a
//     node pulls the contributions of its children, so siblings never write the
//     same parent) and back-substituted in the reverse order;
//   * eliminated rows are recovered in parallel from the solution of the chunk's two
//     ends by the forward rows' recurrence (one worker per chunk again).
// The plan (levels, chunks, heights) depends only on the topology and is built
// once on the host; only values flow per solve.
//
// Layout: every reduction level stores its nodes in "slots", chunk-interleaved:
// node k of chunk c sits at slot k*NC + c (NC = chunks of the level). The workers
// of a wave take consecutive chunks, so each step of a sweep is one coalesced
// load/store of consecutive values instead of a stride-C gather. Levels are
// padded to C*NC slots; padding slots are holes (no next, no chunk ends).
#pragma once
#include <cstdint>
#include <vector>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace nrt {

// Host-side plan.
struct TreePlanHost {
  int C = 32;                       // chunk length of level 0
  int C_hi = 32;                    // chunk length of the levels above
  int levels = 0;                   // reduction levels (L)
  std::vector<int32_t> nslots;      // slots per level, L+1 entries (last = final forest)
  std::vector<int32_t> nchunk;      // chunks per reduction level (L entries, +1 padding)
  std::vector<int32_t> node_off;    // offset of each level in the flat arrays, L+2
  std::vector<int32_t> next;        // flat: slot in level+1 of a surviving node (-1 otherwise)
  std::vector<int32_t> chunk_off;   // per level, L+1 entries, into chunk_len
  std::vector<int32_t> chunk_len;
  std::vector<int32_t> chunk_dyn;   // bit 0: the chunk's matrix values change between
                                    // steps; bit 1: a wide chunk (local mode)
  std::vector<int32_t> slot0;       // level-0 slot of each row (N entries)
  std::vector<int32_t> row0;        // row of each level-0 slot (-1 for padding)
  std::vector<int32_t> fin_par;     // final forest: parent (final-local) or -1
  std::vector<int32_t> fin_order;   // final-forest nodes ordered by height
  std::vector<int32_t> fin_hoff;    // height boundaries into fin_order, H+1 entries
  std::vector<int32_t> fin_cptr, fin_child;  // children of final-forest nodes (CSR)
  int heights = 0;                  // H
  // Chain tasks of the final forest: a node with exactly one child is eliminated
  // right after that child by the same thread (no barrier between), so a task is a
  // maximal single-child chain, bottom to top, and tasks are ordered by task height
  // (1 + the largest task height among the bottom node's children). A forest of
  // segments (start and end node each) takes half as many passes as by heights.
  int ftask_hn = 0;                 // task heights
  std::vector<int32_t> ftask_hoff;  // task-height boundaries into tasks, ftask_hn+1
  std::vector<int32_t> ftask_off;   // task t: nodes [ftask_off[t], ftask_off[t+1])
  std::vector<int32_t> ftask_node;  // ... of ftask_node, bottom to top
  // The same tasks in task positions (index k into ftask_node) for a workgroup that
  // stages the whole forest in LDS at those positions: the position of the parent
  // of each task's top node (-1: a root), and the positions of the children of its
  // bottom node, [ftask_cptr[t], ftask_cptr[t+1]) of ftask_child, in fin_child order.
  std::vector<int32_t> ftask_par, ftask_cptr, ftask_child;
  // ... and as final-local nodes (for the solve from the work arrays): the parent
  // node of each task's top (-1: a root), the children of its bottom node.
  std::vector<int32_t> ftask_pnode, ftask_cnode;
  // Local mode: the chain tasks of each group's local final nodes, group g's tasks
  // of task height h being [ltask_hoff[g*ltask_hn+h], +1) of ltask_off / ltask_node.
  int ltask_hn = 0;
  std::vector<int32_t> ltask_hoff, ltask_off, ltask_node;
  // Level-0 slots of the rows in dynamic chunks (all rows when levels == 0): the
  // rows whose right-hand side a step's update phase writes for a direct solve;
  // the static chunks compute theirs as they are swept (TpsRhsUpdate).
  std::vector<int32_t> dyn0;
  // Final forest split for grid launches (partition_final_forest): the nodes below
  // height fin_cut form whole subtrees, dealt to sub_groups groups that workgroups
  // eliminate side by side; only heights >= fin_cut are left to one workgroup.
  int fin_cut = 0;                  // 0: no split
  int sub_groups = 0;
  std::vector<int32_t> sub_order;   // nodes below fin_cut, by group, then height
  std::vector<int32_t> sub_hoff;    // group g, height h: [sub_hoff[g*cut+h], ..+1)
  // The heights >= fin_cut staged in LDS by the one workgroup that solves them
  // (top_n == 0: not staged). Local entries: the top nodes in height order, then
  // their children below the cut (subtree roots, inputs only).
  int top_n = 0, top_ext = 0;
  std::vector<int32_t> top_node;    // final-local node of each local entry
  std::vector<int32_t> top_hoff;    // height boundaries (heights fin_cut..H-1), local
  std::vector<int32_t> top_cptr, top_child;  // children of top nodes, local indices
  std::vector<int32_t> top_par;     // local parent of each top node, or -1
  // Local mode (build_tree_plan with row groups): G groups of rows, one per
  // workgroup of a grid launch. Chains break where the group changes, so every
  // chunk of every level belongs to one group; a final-forest node is local when
  // its whole subtree is in its group, else it is in the top (closed upwards).
  // Each workgroup reduces, solves its local nodes and back-substitutes alone;
  // only the top needs the whole grid (two grid barriers per solve).
  int groups = 0;
  std::vector<int32_t> gcol;        // level lv, group g: columns [gcol[lv*(G+1)+g], +1)
  std::vector<int32_t> loc_order;   // local final nodes by group, then height
  std::vector<int32_t> loc_hoff;    // group g, height h: [loc_hoff[g*H+h], +1)
  int top_lds_ok = 0;               // the top fits in LDS (grid mode: kTopLdsNodes
                                    // entries; local mode, sized to the top: kTopLdsLocal)
  int top_wide = 0;                 // most top nodes at one height
  int top_narrow = 0;               // first top height (relative) from which no height
                                    // holds more than a wave's 64 nodes
  int top_kids = 0;                 // most children of one top node
  // Group-resident form of local mode: each group's nodes of levels 1..L-1 and its
  // local final nodes get local indices (kept in LDS by the launch). Level lv of
  // group g starts at lbase[g*(L+2)+lv], node k of its column c at
  // + k*nc + (c - c0) (c0, nc: the group's columns of that level); its local final
  // nodes follow at lbase[g*(L+2)+L] in loc_order order; lbase[g*(L+2)+L+1] is
  // the group's count. Chunk-end codes: a local index (>= 0) or -2-u for final
  // node u of the top (kept in the global arrays).
  int lds_nodes = 0;                // largest group count
  int gl_cols = 0, gl_fin = 0, gl_edges = 0;  // largest group: columns, local final nodes, their children
  int gl_wide = 0;                  // largest group: wide chunks
  std::vector<int32_t> wide_off;    // group g, level lv: wide chunks [wide_off[g*(L+1)+lv], +1)
  std::vector<int32_t> wide_col;    // ... their columns (group-relative)
  std::vector<int32_t> lbase;
  std::vector<int32_t> lnext;       // per slot of levels 0..L-1: end code, -1 if no end
  std::vector<int32_t> lf_cptr, lf_child;  // local final nodes (loc_order order): children
  std::vector<int32_t> lf_par;      // parent: local index, top code, or -1
  int64_t total_nodes() const { return node_off.empty() ? 0 : node_off.back(); }
};

// LDS capacity for the staged top of the final forest (local entries).
constexpr int kTopLdsNodes = 1024;
// Local mode keeps a chain of kWideMin+1 .. kWideMax nodes whole above level 0 (a
// wide chunk); the group-resident launch eliminates it with the whole workgroup by
// parallel cyclic reduction, one interior row per thread. Longer chains are cut
// into short chunks as usual, so the next level's chain is shorter: PCR moves
// every row at every step (n log n LDS traffic), the serial sweep n.
constexpr int kWideMin = 64;
constexpr int kWideMax = 2 + 512;
// Local mode stages the top in an LDS region sized to it, up to this many entries
// (56 B each in FP64: 140 KB of the 160 KB per workgroup).
constexpr int kTopLdsLocal = 2560;

// Builds the plan for a forest given by parent indices (par[i] < i or -1), with
// chains cut into chunks of `chunk` rows at level 0 and `chunk_hi` rows above.
// dyn_rows (optional, N entries): rows whose matrix values change between steps;
// chunks holding one, and the reduced chunks that depend on them, are refactored
// every step, the others reuse their factors. Within a level, chunks are placed
// (slot columns) dynamic first, then by decreasing length, so a wave's lanes take
// the same path and stay active together.
// groups (optional, N entries, values 0..G-1): local mode, see TreePlanHost.
// wide: in local mode, keep long chains whole above level 0 (kWideMin).
TreePlanHost build_tree_plan(const std::vector<int32_t>& par, int chunk, int chunk_hi,
                             const std::vector<uint8_t>* dyn_rows = nullptr,
                             const std::vector<int32_t>* groups = nullptr, bool wide = true);

// Splits the final forest of a plan for a launch of `groups` workgroups of
// `threads` threads (any number of workgroups can run it: workgroup b takes the
// groups b, b + nblocks, ...). The cut height minimises a latency model: one unit
// per pass of a workgroup over a height (ceil(nodes / threads)), up and down,
// plus two grid barriers when split; with `stage_top`, heights above a cut whose
// top fits in kTopLdsNodes LDS entries cost a tenth (the lists for that are
// built). No split when groups <= 1.
void partition_final_forest(TreePlanHost& P, int groups, int threads, bool stage_top = false);

// Local mode row groups: G contiguous ranges of rows of (nearly) equal size.
std::vector<int32_t> row_groups(int N, int G);

// Throws when a local-mode plan breaks what its launch relies on: every final node
// local to one group or in the top, the top closed upwards, a local node's children
// in its group, every chunk's columns and ends in its own group.
void check_local_plan(const TreePlanHost& P);

// Parent of each row in an assembled CSR (the single column < row, or -1).
std::vector<int32_t> csr_parents(int N, const int32_t* rowptr, const int32_t* col);

struct TreePlanDev;
// Device view of a plan whose arrays are concatenated in one int32 buffer at
// `base`; with `flat` non-null the concatenation is appended there.
TreePlanDev plan_view(const TreePlanHost& P, const int32_t* base, std::vector<int32_t>* flat);
// Host view with every plan array in an allocation of its own (`store`), each of
// exactly its length: an access past the end of any one array leaves its
// allocation (host emulation under a bounds-checking build).
TreePlanDev plan_view_parts(const TreePlanHost& P, std::vector<std::vector<int32_t>>& store);
constexpr int kPlanParts = 43;

// The matrix with its rows in level-0 slot order (padding slots are empty rows):
// columns as slots, each entry's role (0 diagonal, 1 coupling to the parent, 2 to
// a child), and for HH entry hh the slot of its row and the position of its first
// entry there (hh_mat gives that position in the element-order matrix; the row's
// entries keep their order).
struct SlotCsr {
  std::vector<int32_t> rowptr, col, hh_slot, hh_mat;
  std::vector<double> val;
  std::vector<uint8_t> role;
};
SlotCsr slot_order_csr(const TreePlanHost& P, int N, const int32_t* rowptr, const int32_t* col,
                       const double* val, int H, const int32_t* hh_elem, const int32_t* hh_mat);

// Device view of a plan (pointers into one device allocation).
struct TreePlanDev {
  int levels, heights, ftask_hn, ndyn0, ltask_hn;
  int fin_n, ftask_n, ftask_nc;  // final-forest nodes, chain tasks, bottom-node children
  int fin_cut, sub_groups;
  int top_n, top_ext;
  int groups, top_lds_ok, lds_nodes, top_wide, top_narrow, top_kids;
  int gl_cols, gl_fin, gl_edges, gl_wide;
  const int32_t* nslots;
  const int32_t* nchunk;
  const int32_t* node_off;
  const int32_t* next;
  const int32_t* chunk_off;
  const int32_t* chunk_len;
  const int32_t* chunk_dyn;
  const int32_t* slot0;
  const int32_t* row0;
  const int32_t* fin_par;
  const int32_t* fin_order;
  const int32_t* fin_hoff;
  const int32_t* fin_cptr;
  const int32_t* fin_child;
  const int32_t* ftask_hoff;
  const int32_t* ftask_off;
  const int32_t* ftask_node;
  const int32_t* ftask_par;
  const int32_t* ftask_cptr;
  const int32_t* ftask_child;
  const int32_t* ftask_pnode;
  const int32_t* ftask_cnode;
  const int32_t* dyn0;
  const int32_t* ltask_hoff;
  const int32_t* ltask_off;
  const int32_t* ltask_node;
  const int32_t* sub_order;
  const int32_t* sub_hoff;
  const int32_t* top_node;
  const int32_t* top_hoff;
  const int32_t* top_cptr;
  const int32_t* top_child;
  const int32_t* top_par;
  const int32_t* gcol;
  const int32_t* loc_order;
  const int32_t* loc_hoff;
  const int32_t* lbase;
  const int32_t* lnext;
  const int32_t* lf_cptr;
  const int32_t* lf_child;
  const int32_t* lf_par;
  const int32_t* wide_off;
  const int32_t* wide_col;
};

// Per-value work arrays of one solve, over all levels (total_nodes each): d, up
// (row -> parent coupling), dn (parent -> row coupling), r, x; the elimination
// coefficients ca, cb, cg (interior rows: the forward sweep's r' in ca and fill-in
// g in cb, read by the back-substitution's recurrence; the group-resident form's
// LDS sweeps keep (alpha, beta, gamma) there instead); and the factors kept for the
// steps that reuse them: the forward multipliers f, the pivots' reciprocals and
// the weights of alpha_1 (cw).
template <class T>
struct TreeWork {
  using value_type = T;
  T* d;
  T* up;
  T* dn;
  T* r;
  T* x;
  T* ca;
  T* cb;
  T* cg;
  T* cf;
  T* ci;
  T* cw;  // interior rows: weight of the row's r' in the start row's alpha (WSUM)
};

#if defined(__HIPCC__)
#define NRT_HD __host__ __device__ __forceinline__
#else
#define NRT_HD inline
#endif

// LDS-resident arrays are typed in the LDS address space in device code, so that
// loads and stores through them are LDS instructions, never generic ones (host
// code, e.g. the emulation in tests, uses plain pointers).
#if defined(__HIP_DEVICE_COMPILE__)
#define NRT_LDS __attribute__((address_space(3)))
#else
#define NRT_LDS
#endif
template <class T>
struct TreeWorkL {
  using value_type = T;
  NRT_LDS T* d;
  NRT_LDS T* up;
  NRT_LDS T* dn;
  NRT_LDS T* r;
  NRT_LDS T* x;
  NRT_LDS T* ca;
  NRT_LDS T* cb;
  NRT_LDS T* cg;
  NRT_LDS T* cf;
  NRT_LDS T* ci;
};

// Reciprocal of a pivot. On the device: the hardware estimate refined by Newton
// steps r <- r + r (1 - x r) (two for FP64, one for FP32), within about an ulp and
// a third as many dependent operations as a correctly rounded division, which
// matters on the sweeps' serial paths. On the host: a division. The systems are
// diagonally dominant, so pivots are far from zero, infinity and denormals.
template <class T>
NRT_HD T tps_rcp(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 8) {
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
    return r;
  } else {
    float r = __builtin_amdgcn_rcpf(x);
    r = __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
    return r;
  }
#else
  return T(1) / x;
#endif
}

// Rows of a chunk whose loads are issued together: the sweeps are recurrences, but
// their operands do not depend on them, so a batch of rows is fetched at once and
// the latency is paid once per batch rather than once per row.
constexpr int kTpsBatch = 4;
// Rows per batch of the static sweep and of the back-substitution (their loads do
// not depend on the chain's arithmetic, so a whole batch is one memory round trip).
constexpr int kTpsBatchS = 8;
// Rows per batch of the forward-only static sweep and of the back-substitution from
// r' (more operands per row: the weights, or the recurrence's coefficients).
constexpr int kTpsBatchW = 4;

// Where a chunk's two surviving ends go: the slots of the next level in the same
// arrays (level-synchronous layout). End codes are slot indices.
template <class T>
struct EndsSlots {
  TreeWork<T> w;
  NRT_HD void put(int e, T d, T up, T dn, T r) const {
    w.d[e] = d;
    w.up[e] = up;
    w.dn[e] = dn;
    w.r[e] = r;
  }
  NRT_HD void put_r(int e, T r) const { w.r[e] = r; }
  NRT_HD T x(int e) const { return w.x[e]; }
};

// Right-hand side of a row read from the array r (RowsOf keeps the array's address
// space: LDS arrays of the group-resident form stay LDS accesses).
template <class T>
struct RhsArray {
  const T* r;
  NRT_HD T operator()(int g) const { return r[g]; }
};
template <class Ptr>
struct RowsOf {
  Ptr r;
  NRT_HD auto operator()(int g) const { return r[g]; }
};

// Eliminate the interior of the chunk of m rows at base, base + stride, ... in the
// arrays w; its ends (codes es for row 0, ee for row m-1) go through `ends`.
// WSUM: the start row's reduced right-hand side takes alpha_1 as the sum
// sum_k cw_k r'_k over the interior rows (k = 1, 2, ...), with the weights
// cw_k = (1/pivot_k) prod_{j<k} (-dn_{j+1} / pivot_j) kept in w.cw, so that a
// chunk that reuses its factors needs only the forward sweep (chunk_reduce_static
// <true>) and gets the same value to the bit; the interior rows are then solved
// by the back-substitution's recurrence (chunk_back<MODE, true>).
// rf(g): the right-hand side of row g (the array r, or computed as it is read).
template <bool WSUM, class W, class E, class R>
NRT_HD void chunk_reduce(const W& w, int base, int NC, int m, int es, int ee, const E& ends,
                         const R& rf) {
  using T = typename W::value_type;
  using P = decltype(w.d);
  const P __restrict__ d = w.d;
  const P __restrict__ up = w.up;
  const P __restrict__ dn = w.dn;
  const P __restrict__ ca = w.ca;
  const P __restrict__ cb = w.cb;
  const P __restrict__ cg = w.cg;
  const P __restrict__ cf = w.cf;
  const P __restrict__ ci = w.ci;
  if (m <= 2) {  // nothing to eliminate: the nodes pass through
    for (int k = 0; k < m; k++) {
      const int g = base + k * NC;
      ends.put(k == 0 ? es : ee, d[g], up[g], dn[g], rf(g));
    }
    return;
  }
  // start row s = 0 and first interior row 1
  const T d0 = d[base], u0 = up[base], n0 = dn[base], r0 = rf(base);
  const int g1 = base + NC;
  const T dn1 = dn[g1];
  // forward sweep over rows k = 1..m-1:
  //   d'_k x_k + dn_{k+1} x_{k+1} + g_k x_s = r'_k,
  // stored as (cb, ca) = (g, r') for interior rows (WSUM; else (cb, cg)); row m-1
  // becomes the end row of the reduced system: G_e x_s + D_e x_e + (children) = R_e.
  // Each row's pivot reciprocal is computed once, here, and kept in ci: the
  // backward sweep and the steps that reuse the factors only multiply.
  T dp = d[g1], gk = up[g1], rp = rf(g1);
  cb[g1] = gk;
  if constexpr (WSUM) ca[g1] = rp;  // r' in ca (the back-substitution's recurrence)
  else cg[g1] = rp;
  T pgk = gk, prp = rp;  // values of the row before the last one
  T inv = T(0);          // reciprocal of that row's pivot
  T dnl = dn1;           // dn of the last row
  T wp = T(1), acc = T(0);  // WSUM: running product of the weights, and the sum
  for (int k0 = 2; k0 <= m - 1; k0 += kTpsBatch) {
    T bu[kTpsBatch], bd[kTpsBatch], bn[kTpsBatch], br[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 + j;
      if (k <= m - 1) {
        const int g = base + k * NC;
        bu[j] = up[g];
        bd[j] = d[g];
        bn[j] = dn[g];
        br[j] = rf(g);
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 + j;
      if (k <= m - 1) {
        pgk = gk;
        prp = rp;
        const int g = base + k * NC;
        inv = tps_rcp(dp);
        ci[g - NC] = inv;
        if constexpr (WSUM) {  // row k-1 (interior): its weight, and its share of alpha_1
          const T cwk = inv * wp;
          w.cw[g - NC] = cwk;
          acc = acc + cwk * rp;
          wp = wp * (-bn[j] * inv);
        }
        const T f = bu[j] * inv;
        dp = bd[j] - f * bn[j];
        gk = -f * gk;
        rp = br[j] - f * rp;
        dnl = bn[j];
        cf[g] = f;
        if (k <= m - 2) {
          cb[g] = gk;
          if constexpr (WSUM) ca[g] = rp;
          else cg[g] = rp;
        }
      }
    }
  }
  const T end_d = dp, end_up = gk, end_r = rp;
  if constexpr (WSUM) {
    // backward sweep for beta_1 and gamma_1 only (the reduced start row and the
    // s -> e coupling): the interior keeps (g, r') in (cb, ca), which the
    // back-substitution's recurrence reads (chunk_back<MODE, true>)
    T be = -pgk * inv, ga = -dnl * inv;
    for (int k0 = m - 3; k0 >= 1; k0 -= kTpsBatch) {
      T bi[kTpsBatch], bg[kTpsBatch], bn[kTpsBatch];
#pragma unroll
      for (int j = 0; j < kTpsBatch; j++) {
        const int k = k0 - j;
        if (k >= 1) {
          const int g = base + k * NC;
          bi[j] = ci[g];
          bg[j] = cb[g];
          bn[j] = dn[g + NC];  // coupling k -> k+1
        }
      }
#pragma unroll
      for (int j = 0; j < kTpsBatch; j++) {
        const int k = k0 - j;
        if (k >= 1) {
          const T dcur = bn[j];
          be = (-bg[j] - dcur * be) * bi[j];
          ga = -dcur * ga * bi[j];
        }
      }
    }
    ends.put(es, d0 + dn1 * be, u0, n0, r0 - dn1 * acc);
    ends.put(ee, end_d, end_up, dn1 * ga, end_r);
    return;
  }
  // backward sweep: x_k = alpha_k + beta_k x_s + gamma_k x_e, into (ca, cb, cg)
  T al, be, ga;
  {
    const int g2 = base + (m - 2) * NC;
    al = prp * inv;
    be = -pgk * inv;
    ga = -dnl * inv;
    ca[g2] = al;
    cb[g2] = be;
    cg[g2] = ga;
  }
  for (int k0 = m - 3; k0 >= 1; k0 -= kTpsBatch) {
    T bi[kTpsBatch], bg[kTpsBatch], brr[kTpsBatch], bn[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        const int g = base + k * NC;
        bi[j] = ci[g];
        bg[j] = cb[g];
        brr[j] = cg[g];
        bn[j] = dn[g + NC];  // coupling k -> k+1
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        const T dcur = bn[j];
        al = (brr[j] - dcur * al) * bi[j];
        be = (-bg[j] - dcur * be) * bi[j];
        ga = -dcur * ga * bi[j];
        const int g = base + k * NC;
        ca[g] = al;
        cb[g] = be;
        cg[g] = ga;
      }
    }
  }
  // start row s: up_s x_P + (d_s + dn_1 beta_1) x_s + dn_1 gamma_1 x_e = r_s - dn_1 alpha_1;
  // end row e: its dn is the s -> e coupling
  ends.put(es, d0 + dn1 * be, u0, n0, r0 - dn1 * al);
  ends.put(ee, end_d, end_up, dn1 * ga, end_r);
}
template <bool WSUM = false, class W, class E>
NRT_HD void chunk_reduce(const W& w, int base, int NC, int m, int es, int ee, const E& ends) {
  chunk_reduce<WSUM>(w, base, NC, m, es, ee, ends, RowsOf<decltype(w.r)>{w.r});
}

// Eliminate the interior of chunk `ch` of level `lv` and write its ends to lv+1.
template <class T>
NRT_HD void tps_reduce_chunk(const TreePlanDev& P, const TreeWork<T>& w, int lv, int ch) {
  const int NC = P.nchunk[lv], m = P.chunk_len[ch];
  const int base = P.node_off[lv] + ch - P.chunk_off[lv], nbase = P.node_off[lv + 1];
  const int es = nbase + P.next[base];
  const int ee = m >= 2 ? nbase + P.next[base + (m - 1) * NC] : es;
  chunk_reduce<true>(w, base, NC, m, es, ee, EndsSlots<T>{w});
}

// Right-hand side only, for a chunk whose factors (f, 1/pivot, beta, gamma, the
// reduced matrix values of its ends) are those of its last full elimination: the
// same operations on r as tps_reduce_chunk, in the same order, so the results are
// identical; the sweeps are multiply-adds with no division on the dependent path.
// Right-hand side of a row for the static sweep: stored (RhsArray), or computed
// from the state as a direct solve's update phase would store it (TpsRhsUpdate:
// b = c_m dx V / dT + K dx (+ input while the stimulus is on), the same operations).
template <class T>
struct TpsRhsUpdate {
  const double* x;
  const double* cmdx;
  const double* kdx;
  const double* input;
  double dT;
  bool active;
  NRT_HD T operator()(int g) const {
    double bi = cmdx[g] * x[g] / dT + kdx[g];
    if (active) bi = bi + input[g];
    return (T)bi;
  }
};

template <bool WSUM = false, class W, class E, class R>
NRT_HD void chunk_reduce_static(const W& w, int base, int NC, int m, int es, int ee,
                                const E& ends, const R& r) {
  using T = typename W::value_type;
  using P = decltype(w.d);
  const P __restrict__ ca = w.ca;
  const P __restrict__ dn = w.dn;
  const P __restrict__ cf = w.cf;
  const P __restrict__ ci = w.ci;
  if (m <= 2) {
    for (int k = 0; k < m; k++) ends.put_r(k == 0 ? es : ee, r(base + k * NC));
    return;
  }
  const T r0 = r(base);
  const T dn1 = dn[base + NC];
  if constexpr (WSUM) {
    // forward only: r'_k into ca (the back-substitution solves the interior), alpha_1
    // as chunk_reduce<true> sums it
    const P __restrict__ cw = w.cw;
    T rp = r(base + NC);
    ca[base + NC] = rp;
    T acc = T(0);
    acc = acc + cw[base + NC] * rp;
    for (int k0 = 2; k0 <= m - 1; k0 += kTpsBatchW) {
      T br[kTpsBatchW], bf[kTpsBatchW], bw[kTpsBatchW];
#pragma unroll
      for (int j = 0; j < kTpsBatchW; j++) {
        const int k = k0 + j;
        if (k <= m - 1) {
          const int g = base + k * NC;
          br[j] = r(g);
          bf[j] = cf[g];
          bw[j] = k <= m - 2 ? cw[g] : T(0);
        }
      }
#pragma unroll
      for (int j = 0; j < kTpsBatchW; j++) {
        const int k = k0 + j;
        if (k <= m - 1) {
          rp = br[j] - bf[j] * rp;
          if (k <= m - 2) {
            ca[base + k * NC] = rp;
            acc = acc + bw[j] * rp;
          }
        }
      }
    }
    ends.put_r(ee, rp);
    ends.put_r(es, r0 - dn1 * acc);
    return;
  }
  // forward: rp_k = r_k - f_k rp_{k-1}
  T rp = r(base + NC);
  ca[base + NC] = rp;
  for (int k0 = 2; k0 <= m - 1; k0 += kTpsBatchS) {
    T br[kTpsBatchS], bf[kTpsBatchS];
#pragma unroll
    for (int j = 0; j < kTpsBatchS; j++) {
      const int k = k0 + j;
      if (k <= m - 1) {
        const int g = base + k * NC;
        br[j] = r(g);
        bf[j] = cf[g];
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatchS; j++) {
      const int k = k0 + j;
      if (k <= m - 1) {
        rp = br[j] - bf[j] * rp;
        if (k <= m - 2) ca[base + k * NC] = rp;
      }
    }
  }
  ends.put_r(ee, rp);
  // backward: alpha_k = (rp_k - dn_{k+1} alpha_{k+1}) / pivot_k
  const int g2 = base + (m - 2) * NC;
  T al = ca[g2] * ci[g2];
  ca[g2] = al;
  for (int k0 = m - 3; k0 >= 1; k0 -= kTpsBatchS) {
    T bp[kTpsBatchS], bi[kTpsBatchS], bn[kTpsBatchS];
#pragma unroll
    for (int j = 0; j < kTpsBatchS; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        const int g = base + k * NC;
        bp[j] = ca[g];
        bi[j] = ci[g];
        bn[j] = dn[g + NC];
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatchS; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        al = (bp[j] - bn[j] * al) * bi[j];
        ca[base + k * NC] = al;
      }
    }
  }
  ends.put_r(es, r0 - dn1 * al);
}

template <class W, class E>
NRT_HD void chunk_reduce_static(const W& w, int base, int NC, int m, int es, int ee,
                                const E& ends) {
  chunk_reduce_static(w, base, NC, m, es, ee, ends, RowsOf<decltype(w.r)>{w.r});
}

template <class T, class R>
NRT_HD void tps_reduce_chunk_static(const TreePlanDev& P, const TreeWork<T>& w, int lv, int ch,
                                    const R& rhs) {
  const int NC = P.nchunk[lv], m = P.chunk_len[ch];
  const int base = P.node_off[lv] + ch - P.chunk_off[lv], nbase = P.node_off[lv + 1];
  const int es = nbase + P.next[base];
  const int ee = m >= 2 ? nbase + P.next[base + (m - 1) * NC] : es;
  chunk_reduce_static<true>(w, base, NC, m, es, ee, EndsSlots<T>{w}, rhs);
}
template <class T>
NRT_HD void tps_reduce_chunk_static(const TreePlanDev& P, const TreeWork<T>& w, int lv, int ch) {
  tps_reduce_chunk_static(P, w, lv, ch, RhsArray<T>{w.r});
}

// Final forest, elimination of node v (all of its children are final already).
// The eliminated pivot goes to ca (free at the final level): d must keep the
// reduced matrix value, which a chunk that reuses its factors does not rewrite.
template <class T>
NRT_HD void tps_final_up(const TreePlanDev& P, const TreeWork<T>& w, int v) {
  const int base = P.node_off[P.levels];
  T dv = w.d[base + v], rv = w.r[base + v];
  for (int k = P.fin_cptr[v]; k < P.fin_cptr[v + 1]; k++) {
    const int c = base + P.fin_child[k];
    const T f = w.dn[c] * tps_rcp(w.ca[c]);
    dv = dv - f * w.up[c];
    rv = rv - f * w.r[c];
  }
  w.ca[base + v] = dv;
  w.r[base + v] = rv;
}

template <class T>
NRT_HD void tps_final_down(const TreePlanDev& P, const TreeWork<T>& w, int v) {
  const int base = P.node_off[P.levels];
  const int p = P.fin_par[v];
  T acc = w.r[base + v];
  if (p >= 0) acc = acc - w.up[base + v] * w.x[base + p];
  w.x[base + v] = acc * tps_rcp(w.ca[base + v]);
}

// Final forest by chain tasks (TreePlanHost::ftask_*), task t: its bottom node as
// tps_final_up (children in other tasks), then each node above it from the one
// below, whose pivot and right-hand side are still in registers. The operands that
// do not depend on the chain (the nodes, their d, r, up and dn) are loaded before
// the chain's arithmetic; the arithmetic is tps_final_up's, operand for operand.
constexpr int kTaskBatch = 4;
template <class T>
NRT_HD void tps_task_up(const TreePlanDev& P, const TreeWork<T>& w, const int32_t* toff,
                        const int32_t* tnode, int t) {
  const int base = P.node_off[P.levels];
  const int k0 = toff[t], k1 = toff[t + 1];
  int v = tnode[k0];
  T dv = w.d[base + v], rv = w.r[base + v];
  for (int k = P.fin_cptr[v]; k < P.fin_cptr[v + 1]; k++) {
    const int c = base + P.fin_child[k];
    const T f = w.dn[c] * tps_rcp(w.ca[c]);
    dv = dv - f * w.up[c];
    rv = rv - f * w.r[c];
  }
  w.ca[base + v] = dv;
  w.r[base + v] = rv;
  for (int k = k0 + 1; k < k1; k += kTaskBatch) {
    int u[kTaskBatch];
    T bd[kTaskBatch], br[kTaskBatch], bup[kTaskBatch], bdn[kTaskBatch];
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k + j < k1) {
        const int c = base + tnode[k + j - 1];  // the node below: its only child
        u[j] = base + tnode[k + j];
        bd[j] = w.d[u[j]];
        br[j] = w.r[u[j]];
        bup[j] = w.up[c];
        bdn[j] = w.dn[c];
      }
    }
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k + j < k1) {
        const T f = bdn[j] * tps_rcp(dv);
        dv = bd[j] - f * bup[j];
        rv = br[j] - f * rv;
        w.ca[u[j]] = dv;
        w.r[u[j]] = rv;
      }
    }
  }
}

// Back-substitution of task t, top to bottom (tps_final_down's arithmetic): the
// top node from its parent's solution (a task of a larger height), each node below
// from the one above.
template <class T>
NRT_HD void tps_task_down(const TreePlanDev& P, const TreeWork<T>& w, const int32_t* toff,
                          const int32_t* tnode, int t) {
  const int base = P.node_off[P.levels];
  const int k0 = toff[t], k1 = toff[t + 1];
  const int v = tnode[k1 - 1];
  const int p = P.fin_par[v];
  T acc = w.r[base + v];
  if (p >= 0) acc = acc - w.up[base + v] * w.x[base + p];
  T xv = acc * tps_rcp(w.ca[base + v]);
  w.x[base + v] = xv;
  for (int k = k1 - 2; k >= k0; k -= kTaskBatch) {
    int u[kTaskBatch];
    T br[kTaskBatch], bup[kTaskBatch], bca[kTaskBatch];
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k - j >= k0) {
        u[j] = base + tnode[k - j];
        br[j] = w.r[u[j]];
        bup[j] = w.up[u[j]];
        bca[j] = w.ca[u[j]];
      }
    }
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k - j >= k0) {
        const T a = br[j] - bup[j] * xv;
        xv = a * tps_rcp(bca[j]);
        w.x[u[j]] = xv;
      }
    }
  }
}

// The same chain tasks from the work arrays with the task-indexed lists (ftask_cptr,
// ftask_cnode, ftask_pnode): a task's bounds, child and parent nodes come from
// arrays indexed by the task, so its loads wait on two earlier loads at most rather
// than on a chain of node lookups. Operand for operand tps_task_up/down.
constexpr int kTaskKids = 4;
template <class T>
NRT_HD void tps_ftask_up(const TreePlanDev& P, const TreeWork<T>& w, int t) {
  const int base = P.node_off[P.levels];
  const int k0 = P.ftask_off[t], k1 = P.ftask_off[t + 1];
  const int q0 = P.ftask_cptr[t], q1 = P.ftask_cptr[t + 1];
  const int v = base + P.ftask_node[k0];
  T dv = w.d[v], rv = w.r[v];
  for (int q = q0; q < q1; q += kTaskKids) {
    int c[kTaskKids];
    T bdn[kTaskKids], bca[kTaskKids], bup[kTaskKids], br[kTaskKids];
#pragma unroll
    for (int j = 0; j < kTaskKids; j++)
      if (q + j < q1) c[j] = base + P.ftask_cnode[q + j];
#pragma unroll
    for (int j = 0; j < kTaskKids; j++) {
      if (q + j < q1) {
        bdn[j] = w.dn[c[j]];
        bca[j] = w.ca[c[j]];
        bup[j] = w.up[c[j]];
        br[j] = w.r[c[j]];
      }
    }
#pragma unroll
    for (int j = 0; j < kTaskKids; j++) {
      if (q + j < q1) {
        const T f = bdn[j] * tps_rcp(bca[j]);
        dv = dv - f * bup[j];
        rv = rv - f * br[j];
      }
    }
  }
  w.ca[v] = dv;
  w.r[v] = rv;
  for (int k = k0 + 1; k < k1; k += kTaskBatch) {
    int u[kTaskBatch];
    T bd[kTaskBatch], br[kTaskBatch], bup[kTaskBatch], bdn[kTaskBatch];
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k + j < k1) {
        const int c = base + P.ftask_node[k + j - 1];  // the node below: its only child
        u[j] = base + P.ftask_node[k + j];
        bd[j] = w.d[u[j]];
        br[j] = w.r[u[j]];
        bup[j] = w.up[c];
        bdn[j] = w.dn[c];
      }
    }
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k + j < k1) {
        const T f = bdn[j] * tps_rcp(dv);
        dv = bd[j] - f * bup[j];
        rv = br[j] - f * rv;
        w.ca[u[j]] = dv;
        w.r[u[j]] = rv;
      }
    }
  }
}
template <class T>
NRT_HD void tps_ftask_down(const TreePlanDev& P, const TreeWork<T>& w, int t) {
  const int base = P.node_off[P.levels];
  const int k0 = P.ftask_off[t], k1 = P.ftask_off[t + 1];
  const int pn = P.ftask_pnode[t];
  const int v = base + P.ftask_node[k1 - 1];
  T acc = w.r[v];
  if (pn >= 0) acc = acc - w.up[v] * w.x[base + pn];
  T xv = acc * tps_rcp(w.ca[v]);
  w.x[v] = xv;
  for (int k = k1 - 2; k >= k0; k -= kTaskBatch) {
    int u[kTaskBatch];
    T br[kTaskBatch], bup[kTaskBatch], bca[kTaskBatch];
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k - j >= k0) {
        u[j] = base + P.ftask_node[k - j];
        br[j] = w.r[u[j]];
        bup[j] = w.up[u[j]];
        bca[j] = w.ca[u[j]];
      }
    }
#pragma unroll
    for (int j = 0; j < kTaskBatch; j++) {
      if (k - j >= k0) {
        const T a = br[j] - bup[j] * xv;
        xv = a * tps_rcp(bca[j]);
        w.x[u[j]] = xv;
      }
    }
  }
}

// The whole final forest staged in LDS by the one workgroup that solves it (block
// launches), at task positions: one load of each node's d, r, up, dn and of the
// task lists, every task height up and down in LDS with a workgroup barrier each,
// one store of x. The arithmetic is tps_task_up/down's, operand for operand, so
// the solution is the same to the bit. The eliminated pivot overwrites d and the
// solution overwrites r (neither is needed again), and the global d, r and ca of
// the final nodes are left as they were.
template <class T>
struct ForestLds {
  T* d;
  T* r;
  T* up;
  T* dn;
  int32_t* node;   // final-local node of each position
  int32_t* toff;   // task t: positions [toff[t], toff[t+1])
  int32_t* cptr;   // children of task t's bottom node: [cptr[t], cptr[t+1]) of child
  int32_t* child;
  int32_t* par;    // position of the parent of task t's top node, or -1
  int32_t* thoff;  // task height h: tasks [thoff[h], thoff[h+1])
  NRT_HD static ForestLds carve(void* p, const TreePlanDev& P) {
    const int n = P.fin_n, nt = P.ftask_n;
    ForestLds L;
    L.d = (T*)p;
    L.r = L.d + n;
    L.up = L.r + n;
    L.dn = L.up + n;
    L.node = (int32_t*)(L.dn + n);
    L.toff = L.node + n;
    L.cptr = L.toff + nt + 1;
    L.child = L.cptr + nt + 1;
    L.par = L.child + P.ftask_nc;
    L.thoff = L.par + nt;
    return L;
  }
};
// LDS bytes of a staged final forest (0: the plan has no chain tasks).
template <class T>
NRT_HD size_t forest_lds_bytes(const TreePlanDev& P) {
  if (P.ftask_hn <= 0) return 0;
  const size_t n = (size_t)P.fin_n, nt = (size_t)P.ftask_n;
  const size_t b = 4 * n * sizeof(T) +
                   (n + 2 * (nt + 1) + (size_t)P.ftask_nc + nt + (size_t)P.ftask_hn + 1) *
                       sizeof(int32_t);
  return (b + 15) & ~(size_t)15;
}
// Staging, lane j < forest_stage_count: node values and task lists.
template <class T>
NRT_HD void forest_stage(const TreePlanDev& P, const TreeWork<T>& w, const ForestLds<T>& L,
                         int j) {
  if (j < P.fin_n) {
    const int v = P.ftask_node[j];
    const int g = P.node_off[P.levels] + v;
    L.node[j] = v;
    L.d[j] = w.d[g];
    L.r[j] = w.r[g];
    L.up[j] = w.up[g];
    L.dn[j] = w.dn[g];
  }
  if (j <= P.ftask_n) {
    L.toff[j] = P.ftask_off[j];
    L.cptr[j] = P.ftask_cptr[j];
    if (j < P.ftask_n) L.par[j] = P.ftask_par[j];
  }
  if (j < P.ftask_nc) L.child[j] = P.ftask_child[j];
  if (j <= P.ftask_hn) L.thoff[j] = P.ftask_hoff[j];
}
NRT_HD int forest_stage_count(const TreePlanDev& P) {
  const int a = P.fin_n > P.ftask_n + 1 ? P.fin_n : P.ftask_n + 1;
  const int b = P.ftask_nc > P.ftask_hn + 1 ? P.ftask_nc : P.ftask_hn + 1;
  return a > b ? a : b;
}
// Task t, bottom to top (tps_task_up).
template <class T>
NRT_HD void forest_up(const ForestLds<T>& L, int t) {
  const int k0 = L.toff[t], k1 = L.toff[t + 1];
  T dv = L.d[k0], rv = L.r[k0];
  for (int q = L.cptr[t]; q < L.cptr[t + 1]; q++) {
    const int c = L.child[q];
    const T f = L.dn[c] * tps_rcp(L.d[c]);
    dv = dv - f * L.up[c];
    rv = rv - f * L.r[c];
  }
  L.d[k0] = dv;
  L.r[k0] = rv;
  for (int k = k0 + 1; k < k1; k++) {
    const T f = L.dn[k - 1] * tps_rcp(dv);
    dv = L.d[k] - f * L.up[k - 1];
    rv = L.r[k] - f * rv;
    L.d[k] = dv;
    L.r[k] = rv;
  }
}
// Task t, top to bottom (tps_task_down): the solution replaces r.
template <class T>
NRT_HD void forest_down(const ForestLds<T>& L, int t) {
  const int k0 = L.toff[t], k1 = L.toff[t + 1];
  const int p = L.par[t];
  T acc = L.r[k1 - 1];
  if (p >= 0) acc = acc - L.up[k1 - 1] * L.r[p];
  T xv = acc * tps_rcp(L.d[k1 - 1]);
  L.r[k1 - 1] = xv;
  for (int k = k1 - 2; k >= k0; k--) {
    const T a = L.r[k] - L.up[k] * xv;
    xv = a * tps_rcp(L.d[k]);
    L.r[k] = xv;
  }
}
template <class T>
NRT_HD void forest_store(const TreePlanDev& P, const TreeWork<T>& w, const ForestLds<T>& L,
                         int j) {
  w.x[P.node_off[P.levels] + L.node[j]] = L.r[j];
}

// The top of the final forest staged in LDS (TreePlanHost::top_*): one workgroup
// loads the top nodes and their children below the cut once, eliminates and
// back-substitutes every top height there, and writes back only x. The arithmetic
// is tps_final_up/down's, operand for operand.
template <class T>
struct TopLds {
  T* d;  // d, then the pivot after elimination (children below the cut: their pivot)
  T* r;
  T* up;
  T* dn;
  T* x;
  int32_t* cptr;
  int32_t* child;
  int32_t* par;
  int32_t* hoff;
  // carve from one buffer of top_lds_bytes<T>(n) for n >= top_n + top_ext entries
  NRT_HD static TopLds carve(void* p, int n = kTopLdsNodes) {
    T* t = (T*)p;
    int32_t* i = (int32_t*)(t + 5 * n);
    return TopLds{t, t + n, t + 2 * n, t + 3 * n, t + 4 * n, i, i + n + 1, i + 2 * n + 1,
                  i + 3 * n + 1};
  }
};
template <class T>
constexpr size_t top_lds_bytes(int n = kTopLdsNodes) {
  return 5 * (size_t)n * sizeof(T) + (4 * (size_t)n + 2) * sizeof(int32_t);
}

// Index j of every staged list: the values of entry j (j < top_n + top_ext) and
// the plan lists' j-th items; run for j < top_stage_count(P).
NRT_HD int top_stage_count(const TreePlanDev& P) {
  const int ne = P.top_n + P.top_ext, nh = P.heights - P.fin_cut;
  return ne > nh + 1 ? ne : nh + 1;
}
template <class T>
NRT_HD void top_stage(const TreePlanDev& P, const TreeWork<T>& w, const TopLds<T>& L, int j) {
  if (j < P.top_n + P.top_ext) {
    const int g = P.node_off[P.levels] + P.top_node[j];
    L.d[j] = j < P.top_n ? w.d[g] : w.ca[g];
    L.r[j] = w.r[g];
    L.up[j] = w.up[g];
    L.dn[j] = w.dn[g];
  }
  if (j < P.top_n) {
    L.par[j] = P.top_par[j];
    L.cptr[j + 1] = P.top_cptr[j + 1];
  }
  if (j == 0) L.cptr[0] = P.top_cptr[0];
  if (j < P.top_cptr[P.top_n]) L.child[j] = P.top_child[j];
  if (j <= P.heights - P.fin_cut) L.hoff[j] = P.top_hoff[j];
}

// Values only (the lists of a launch are fixed): entry j < top_n + top_ext.
template <class T>
NRT_HD void top_stage_values(const TreePlanDev& P, const TreeWork<T>& w, const TopLds<T>& L,
                             int j) {
  const int g = P.node_off[P.levels] + P.top_node[j];
  L.d[j] = j < P.top_n ? w.d[g] : w.ca[g];
  L.r[j] = w.r[g];
  L.up[j] = w.up[g];
  L.dn[j] = w.dn[g];
}

template <class T>
NRT_HD void top_up(const TopLds<T>& L, int j) {
  T dv = L.d[j], rv = L.r[j];
  for (int k = L.cptr[j]; k < L.cptr[j + 1]; k++) {
    const int c = L.child[k];
    const T f = L.dn[c] * tps_rcp(L.d[c]);
    dv = dv - f * L.up[c];
    rv = rv - f * L.r[c];
  }
  L.d[j] = dv;
  L.r[j] = rv;
}

template <class T>
NRT_HD void top_down(const TopLds<T>& L, int j) {
  const int p = L.par[j];
  T acc = L.r[j];
  if (p >= 0) acc = acc - L.up[j] * L.x[p];
  L.x[j] = acc * tps_rcp(L.d[j]);
}

// After the down sweep: x of top node j back to the final level.
template <class T>
NRT_HD void top_store(const TreePlanDev& P, const TreeWork<T>& w, const TopLds<T>& L, int j) {
  w.x[P.node_off[P.levels] + P.top_node[j]] = L.x[j];
}

// Stores of the back-substitution: into the level's x (OUT_LEVEL), or at level 0
// into the caller's FP64 vector, which is in level-0 slot order (level 0 starts at
// offset 0), assigned or accumulated (a refinement).
enum TpsOut { OUT_LEVEL = 0, OUT_ASSIGN = 1, OUT_ADD = 2 };

template <int MODE, class W, class T>
NRT_HD void tps_put(const W& w, int g, T v, double* out) {
  if (MODE == OUT_LEVEL) w.x[g] = v;
  else out[g] = MODE == OUT_ADD ? out[g] + (double)v : (double)v;
}

// Back-substitution of the chunk of m rows at base, base + NC, ... in w from the
// solution at its ends (codes es, ee, read through `ends`).
// FROM_RP: chunks swept by chunk_reduce<true> or chunk_reduce_static<true> (ca holds
// r'_k, cb the fill-in g_k): the interior by the back-substitution's recurrence.
template <int MODE, bool FROM_RP = false, class W, class E>
NRT_HD void chunk_back(const W& w, int base, int NC, int m, int es, int ee, const E& ends,
                       double* out) {
  using T = typename W::value_type;
  using P = decltype(w.d);
  const int ge = base + (m - 1) * NC;
  const T xs = ends.x(es);
  const T xe = m >= 2 ? ends.x(ee) : xs;
  tps_put<MODE>(w, base, xs, out);
  if (m >= 2) tps_put<MODE>(w, ge, xe, out);
  const P __restrict__ ca = w.ca;
  const P __restrict__ cb = w.cb;
  const P __restrict__ cg = w.cg;
  if constexpr (FROM_RP) {
    // the forward sweep's rows d'_k x_k + dn_{k+1} x_{k+1} + g_k x_s = r'_k, solved
    // from the end down: x_k = (r'_k - dn_{k+1} x_{k+1} - g_k x_s) / d'_k, with
    // (r', g) in (ca, cb)
    const P __restrict__ ci = w.ci;
    const P __restrict__ dn = w.dn;
    T xn = xe;
    for (int k0 = m - 2; k0 >= 1; k0 -= kTpsBatchW) {
      T bp[kTpsBatchW], bi[kTpsBatchW], bn[kTpsBatchW], bg[kTpsBatchW];
#pragma unroll
      for (int j = 0; j < kTpsBatchW; j++) {
        const int k = k0 - j;
        if (k >= 1) {
          const int gg = base + k * NC;
          bp[j] = ca[gg];
          bi[j] = ci[gg];
          bn[j] = dn[gg + NC];
          bg[j] = cb[gg];
        }
      }
#pragma unroll
      for (int j = 0; j < kTpsBatchW; j++) {
        const int k = k0 - j;
        if (k >= 1) {
          xn = (bp[j] - bn[j] * xn - bg[j] * xs) * bi[j];
          tps_put<MODE>(w, base + k * NC, xn, out);
        }
      }
    }
    return;
  }
  for (int k0 = 1; k0 <= m - 2; k0 += kTpsBatchS) {
    T a[kTpsBatchS], b[kTpsBatchS], g[kTpsBatchS];
#pragma unroll
    for (int j = 0; j < kTpsBatchS; j++) {
      const int k = k0 + j;
      if (k <= m - 2) {
        const int gg = base + k * NC;
        a[j] = ca[gg];
        b[j] = cb[gg];
        g[j] = cg[gg];
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatchS; j++) {
      const int k = k0 + j;
      if (k <= m - 2) tps_put<MODE>(w, base + k * NC, a[j] + b[j] * xs + g[j] * xe, out);
    }
  }
}

// Back-substitution of chunk `ch` of level `lv` from the solution of level lv+1
// (after tps_reduce_chunk or tps_reduce_chunk_static).
template <int MODE, class T>
NRT_HD void tps_back_chunk(const TreePlanDev& P, const TreeWork<T>& w, int lv, int ch,
                           double* out) {
  const int NC = P.nchunk[lv], m = P.chunk_len[ch];
  const int base = P.node_off[lv] + ch - P.chunk_off[lv], nbase = P.node_off[lv + 1];
  const int es = nbase + P.next[base];
  const int ee = m >= 2 ? nbase + P.next[base + (m - 1) * NC] : es;
  chunk_back<MODE, true>(w, base, NC, m, es, ee, EndsSlots<T>{w}, out);
}

// Level 0 as the final forest itself (a system with no chain longer than 2).
template <int MODE, class T>
NRT_HD void tps_final_out(const TreePlanDev& P, const TreeWork<T>& w, int v, double* out) {
  (void)P;
  tps_put<MODE>(w, v, w.x[v], out);
}

// ---- group-resident local mode (TreePlanHost::lbase): level 0 and the top in the
// global arrays w, every other node of group g in its local arrays lw. The
// arithmetic is the level-synchronous routines', operand for operand.
// Values of top nodes go to wt (and wt2 as well when both: the factor phase fills
// both of a launch's alternating top-input buffers); the top's solution is read
// from w.x.
template <class T>
struct EndsLocal {
  TreeWorkL<T> lw;
  TreeWork<T> w, wt, wt2;
  int fbase;  // first final-forest slot
  bool both;
  NRT_HD void put(int e, T d, T up, T dn, T r) const {
    if (e >= 0) {
      lw.d[e] = d;
      lw.up[e] = up;
      lw.dn[e] = dn;
      lw.r[e] = r;
    } else {
      const int t = fbase - 2 - e;
      wt.d[t] = d;
      wt.up[t] = up;
      wt.dn[t] = dn;
      wt.r[t] = r;
      if (both) {
        wt2.d[t] = d;
        wt2.up[t] = up;
        wt2.dn[t] = dn;
        wt2.r[t] = r;
      }
    }
  }
  NRT_HD void put_r(int e, T r) const {
    if (e >= 0) {
      lw.r[e] = r;
    } else {
      wt.r[fbase - 2 - e] = r;
      if (both) wt2.r[fbase - 2 - e] = r;
    }
  }
  NRT_HD T x(int e) const { return e >= 0 ? lw.x[e] : w.x[fbase - 2 - e]; }
};

// Global arrays of the group-resident form: level 0 and the top's solution (w),
// the top inputs of this solve (wt; wt2 too when both).
template <class T>
struct GlIO {
  TreeWork<T> w, wt, wt2;
  bool both;
  NRT_HD EndsLocal<T> ends(const TreeWorkL<T>& lw, int fbase) const {
    return EndsLocal<T>{lw, w, wt, wt2, fbase, both};
  }
};

// Plan metadata of one group in the group-resident form, staged once per launch
// (LDS on the device) so that no step waits on a chain of plan loads:
//   lvl[7*lv ..]: first column, column count, local base, first col entry, NC of
//                 level lv, first and count of its wide chunks in wcol
//   col[4*q ..]: chunk length, start code, end code, dyn flag of the group's column q
//   scal: first final-forest slot in w, local base of the final nodes, final count
//   hoff: the local final nodes' height boundaries (group-relative, H+1)
//   fpar, fu, fcp, fch: their parent codes, final ids, children (local indices)
//   wcol: the group's wide chunks (columns), level by level
struct GlMeta {
  NRT_LDS int32_t* lvl;
  NRT_LDS int32_t* col;
  NRT_LDS int32_t* scal;
  NRT_LDS int32_t* hoff;
  NRT_LDS int32_t* fpar;
  NRT_LDS int32_t* fu;
  NRT_LDS int32_t* fcp;
  NRT_LDS int32_t* fch;
  NRT_LDS int32_t* wcol;
};

// Byte layout of a group-resident launch's dynamic LDS: the staged top (when
// used), the group's value arrays, its metadata.
struct GlLayout {
  size_t top, values, meta, total;
};
NRT_HD size_t gl_round16(size_t b) { return (b + 15) / 16 * 16; }
NRT_HD int gl_top_entries(const TreePlanDev& P) {
  const int n = P.top_n + P.top_ext;
  return n > 0 ? n : 1;
}
NRT_HD size_t gl_meta_ints(const TreePlanDev& P) {
  return (size_t)7 * P.levels + 4 + (P.heights + 1) + 4 * (size_t)P.gl_cols + 3 * P.gl_fin + 1 +
         P.gl_edges + P.gl_wide;
}
template <class T>
NRT_HD GlLayout gl_layout(const TreePlanDev& P, bool top) {
  GlLayout o;
  o.top = top ? gl_round16(top_lds_bytes<T>(gl_top_entries(P))) : 0;
  o.values = gl_round16((size_t)10 * P.lds_nodes * sizeof(T));
  o.meta = gl_round16(gl_meta_ints(P) * sizeof(int32_t));
  o.total = o.top + o.values + o.meta;
  return o;
}
NRT_HD GlMeta gl_meta_carve(void* p, const TreePlanDev& P) {
  NRT_LDS int32_t* i = (NRT_LDS int32_t*)p;
  GlMeta M;
  M.lvl = i;
  M.col = M.lvl + 7 * P.levels;
  M.scal = M.col + 4 * P.gl_cols;
  M.hoff = M.scal + 4;
  M.fpar = M.hoff + P.heights + 1;
  M.fu = M.fpar + P.gl_fin;
  M.fcp = M.fu + P.gl_fin;
  M.fch = M.fcp + P.gl_fin + 1;
  M.wcol = M.fch + P.gl_edges;
  return M;
}

// Stage group g's metadata; workers tid of nt (all entries once the caller's
// workers have run and synchronised).
NRT_HD void gl_meta_stage(const TreePlanDev& P, int g, const GlMeta& M, int tid, int nt) {
  const int G = P.groups, L = P.levels, H = P.heights;
  int mb = 0;
  for (int lv = 0; lv < L; lv++) {
    const int c0 = P.gcol[lv * (G + 1) + g], nc = P.gcol[lv * (G + 1) + g + 1] - c0;
    const int NC = P.nchunk[lv], slot0 = P.node_off[lv];
    const int w0 = P.wide_off[g * (L + 1)], wb = P.wide_off[g * (L + 1) + lv];
    const int wn = P.wide_off[g * (L + 1) + lv + 1] - wb;
    if (tid == 0) {
      M.lvl[7 * lv] = c0;
      M.lvl[7 * lv + 1] = nc;
      M.lvl[7 * lv + 2] = lv > 0 ? P.lbase[g * (L + 2) + lv] : 0;
      M.lvl[7 * lv + 3] = mb;
      M.lvl[7 * lv + 4] = NC;
      M.lvl[7 * lv + 5] = wb - w0;
      M.lvl[7 * lv + 6] = wn;
    }
    for (int i = tid; i < wn; i += nt) M.wcol[wb - w0 + i] = P.wide_col[wb + i];
    for (int j = tid; j < nc; j += nt) {
      const int ch = P.chunk_off[lv] + c0 + j, m = P.chunk_len[ch];
      const int slot = slot0 + c0 + j;
      M.col[4 * (mb + j)] = m;
      M.col[4 * (mb + j) + 1] = P.lnext[slot];
      M.col[4 * (mb + j) + 2] = m >= 2 ? P.lnext[slot + (m - 1) * NC] : P.lnext[slot];
      M.col[4 * (mb + j) + 3] = P.chunk_dyn[ch];
    }
    mb += nc;
  }
  const int k0 = P.loc_hoff[g * H], nf = P.loc_hoff[g * H + H] - k0;
  if (tid == 0) {
    M.scal[0] = P.node_off[L];
    M.scal[1] = P.lbase[g * (L + 2) + L];
    M.scal[2] = nf;
    M.scal[3] = 0;
  }
  for (int h = tid; h <= H; h += nt) M.hoff[h] = P.loc_hoff[g * H + h] - k0;
  const int e0 = P.lf_cptr[k0];
  for (int j = tid; j < nf; j += nt) {
    M.fpar[j] = P.lf_par[k0 + j];
    M.fu[j] = P.loc_order[k0 + j];
    M.fcp[j] = P.lf_cptr[k0 + j] - e0;
  }
  if (tid == 0) M.fcp[nf] = P.lf_cptr[k0 + nf] - e0;
  for (int q = tid; q < P.lf_cptr[k0 + nf] - e0; q += nt) M.fch[q] = P.lf_child[e0 + q];
}

// Group column j of level lv from the staged metadata (level 0 rows in w, the
// rows above in lw): eliminated in full (how == kGlFull), in full when its values
// change between steps (kGlStep), or with its kept factors (kGlStatic).
enum { kGlFull = 0, kGlStep = 1, kGlStatic = 2 };
template <bool L0, class T>
NRT_HD void glm_reduce(const GlIO<T>& io, const TreeWorkL<T>& lw, const GlMeta& M, int lv,
                       int j, int how) {
  const TreeWork<T>& w = io.w;
  const int c0 = M.lvl[7 * lv], nc = M.lvl[7 * lv + 1];
  const int q = 4 * (M.lvl[7 * lv + 3] + j);
  const int m = M.col[q], es = M.col[q + 1], ee = M.col[q + 2];
  if (M.col[q + 3] & 2) return;  // a wide chunk: the workgroup's (wide_reduce)
  const bool full = how == kGlFull || (how == kGlStep && (M.col[q + 3] & 1));
  const EndsLocal<T> E = io.ends(lw, M.scal[0]);
  if (L0) {
    const int NC = M.lvl[7 * lv + 4];
    if (full) chunk_reduce(w, c0 + j, NC, m, es, ee, E);
    else chunk_reduce_static(w, c0 + j, NC, m, es, ee, E);
  } else {
    const int base = M.lvl[7 * lv + 2] + j;
    if (full) chunk_reduce(lw, base, nc, m, es, ee, E);
    else chunk_reduce_static(lw, base, nc, m, es, ee, E);
  }
}

template <bool L0, int MODE, class T>
NRT_HD void glm_back(const GlIO<T>& io, const TreeWorkL<T>& lw, const GlMeta& M, int lv,
                     int j, double* out) {
  const TreeWork<T>& w = io.w;
  const int c0 = M.lvl[7 * lv], nc = M.lvl[7 * lv + 1];
  const int q = 4 * (M.lvl[7 * lv + 3] + j);
  const int m = M.col[q], es = M.col[q + 1], ee = M.col[q + 2];
  if (M.col[q + 3] & 2) return;  // a wide chunk: the workgroup's (wide_back)
  const EndsLocal<T> E = io.ends(lw, M.scal[0]);
  if (L0) chunk_back<MODE>(w, c0 + j, M.lvl[7 * lv + 4], m, es, ee, E, out);
  else chunk_back<OUT_LEVEL>(lw, M.lvl[7 * lv + 2] + j, nc, m, es, ee, E, out);
}

// Local final node j of the group (group-relative, in height order).
template <class T>
NRT_HD void glm_final_up(const GlIO<T>& io, const TreeWorkL<T>& lw, const GlMeta& M, int j) {
  const TreeWork<T>& w = io.wt;
  const int li = M.scal[1] + j;
  T dv = lw.d[li], rv = lw.r[li];
  for (int q = M.fcp[j]; q < M.fcp[j + 1]; q++) {
    const int c = M.fch[q];
    const T f = lw.dn[c] * tps_rcp(lw.ca[c]);
    dv = dv - f * lw.up[c];
    rv = rv - f * lw.r[c];
  }
  lw.ca[li] = dv;
  lw.r[li] = rv;
  if (M.fpar[j] <= -2) {  // under the top: hand over to the global arrays
    const int t = M.scal[0] + M.fu[j];
    w.ca[t] = dv;
    w.r[t] = rv;
    w.up[t] = lw.up[li];
    w.dn[t] = lw.dn[li];
  }
}

template <class T>
NRT_HD void glm_final_down(const GlIO<T>& io, const TreeWorkL<T>& lw, const GlMeta& M, int j) {
  const TreeWork<T>& w = io.w;
  const int li = M.scal[1] + j, p = M.fpar[j];
  T acc = lw.r[li];
  if (p != -1) {
    const T xp = p >= 0 ? lw.x[p] : w.x[M.scal[0] - 2 - p];
    acc = acc - lw.up[li] * xp;
  }
  lw.x[li] = acc * tps_rcp(lw.ca[li]);
}

// ---- wide chunks: parallel cyclic reduction (PCR) of a chain's interior.
// The m rows of the chunk are at base + k*st in lw (k = 0 .. m-1); the interior
// rows 1 .. m-2 are the unknowns, rows 0 and m-1 (s, e) parameters, as for
// chunk_reduce: x_k = alpha_k + beta_k x_s + gamma_k x_e. Row k reads
// a_k x_{k-1} + d_k x_k + c_k x_{k+1} = r_k with a_k = up[k], c_k = dn[k+1]. Three
// right-hand sides go through the reduction: r, rb (x_s's coefficient moved to
// the right: -a_1 at row 1) and rg (-c_{m-2} at row m-2). Working arrays: a in
// cf, c in ci, d in x, r in r (a per-solve value), rb in cb, rg in cg; the
// matrix arrays (d, up, dn) are left as they are. Each step combines row i with
// rows i - h and i + h (h = 1, 2, 4, ...) until no coupling is left; then
// alpha = r/d, beta = rb/d, gamma = rg/d land in ca, cb, cg.
template <class T>
struct WideRow {
  T a, d, c, r, rb, rg;
};

template <class T>
NRT_HD void wide_init(const TreeWorkL<T>& lw, int base, int st, int m, int k) {
  const int g = base + k * st;
  lw.cf[g] = k == 1 ? T(0) : lw.up[g];
  lw.ci[g] = k == m - 2 ? T(0) : lw.dn[g + st];
  lw.x[g] = lw.d[g];
  lw.cb[g] = k == 1 ? -lw.up[g] : T(0);
  lw.cg[g] = k == m - 2 ? -lw.dn[g + st] : T(0);
}

// New values of interior row k at stride h (read phase of a step).
template <class T>
NRT_HD WideRow<T> wide_step(const TreeWorkL<T>& lw, int base, int st, int m, int k, int h) {
  const int g = base + k * st;
  WideRow<T> o{lw.cf[g], lw.x[g], lw.ci[g], lw.r[g], lw.cb[g], lw.cg[g]};
  if (k - h >= 1) {
    const int q = g - h * st;
    const T k1 = o.a * tps_rcp(lw.x[q]);
    o.d = o.d - k1 * lw.ci[q];
    o.r = o.r - k1 * lw.r[q];
    o.rb = o.rb - k1 * lw.cb[q];
    o.rg = o.rg - k1 * lw.cg[q];
    o.a = -k1 * lw.cf[q];
  } else {
    o.a = T(0);
  }
  if (k + h <= m - 2) {
    const int q = g + h * st;
    const T k2 = o.c * tps_rcp(lw.x[q]);
    o.d = o.d - k2 * lw.cf[q];
    o.r = o.r - k2 * lw.r[q];
    o.rb = o.rb - k2 * lw.cb[q];
    o.rg = o.rg - k2 * lw.cg[q];
    o.c = -k2 * lw.ci[q];
  } else {
    o.c = T(0);
  }
  return o;
}

template <class T>
NRT_HD void wide_put(const TreeWorkL<T>& lw, int base, int st, int k, const WideRow<T>& o) {
  const int g = base + k * st;
  lw.cf[g] = o.a;
  lw.x[g] = o.d;
  lw.ci[g] = o.c;
  lw.r[g] = o.r;
  lw.cb[g] = o.rb;
  lw.cg[g] = o.rg;
}

// After the last step: alpha, beta, gamma of interior row k.
template <class T>
NRT_HD void wide_finish(const TreeWorkL<T>& lw, int base, int st, int k) {
  const int g = base + k * st;
  const T inv = tps_rcp(lw.x[g]);
  lw.ca[g] = lw.r[g] * inv;
  lw.cb[g] = lw.cb[g] * inv;
  lw.cg[g] = lw.cg[g] * inv;
}

// The two ends of the reduced chunk (one worker): as chunk_reduce's, from the
// interior's alpha, beta, gamma next to them. r of rows 1 and m-2 were reduced in
// place, so the end rows take r from the row's own (untouched) slots.
template <class T, class E>
NRT_HD void wide_ends(const TreeWorkL<T>& lw, int base, int st, int m, int es, int ee,
                      const E& ends) {
  const int g0 = base, g1 = base + st, gl = base + (m - 2) * st, ge = base + (m - 1) * st;
  const T dn1 = lw.dn[g1], ae = lw.up[ge];
  ends.put(es, lw.d[g0] + dn1 * lw.cb[g1], lw.up[g0], lw.dn[g0], lw.r[g0] - dn1 * lw.ca[g1]);
  ends.put(ee, lw.d[ge] + ae * lw.cg[gl], ae * lw.cb[gl], dn1 * lw.cg[g1],
           lw.r[ge] - ae * lw.ca[gl]);
}

// Back-substitution of a wide chunk from its ends' solution, rows k = k0, k0+dk, ...
template <class T>
NRT_HD void wide_back_rows(const TreeWorkL<T>& lw, int base, int st, int m, T xs, T xe, int k0,
                           int dk) {
  for (int k = k0; k < m; k += dk) {
    const int g = base + k * st;
    lw.x[g] = k == 0 ? xs : (k == m - 1 ? xe : lw.ca[g] + lw.cb[g] * xs + lw.cg[g] * xe);
  }
}

// Steps of a PCR over n interior rows: strides 1, 2, 4, ... below n.
NRT_HD int wide_steps(int n) {
  int s = 0;
  for (int h = 1; h < n; h *= 2) s++;
  return s;
}

// Local arrays of n entries each, carved from one buffer (LDS on the device).
template <class T>
NRT_HD TreeWorkL<T> gl_carve(void* p, int n) {
  NRT_LDS T* t = (NRT_LDS T*)p;
  return TreeWorkL<T>{t,         t + n,     t + 2 * n, t + 3 * n, t + 4 * n,
                      t + 5 * n, t + 6 * n, t + 7 * n, t + 8 * n, t + 9 * n};
}

}  // namespace nrt
