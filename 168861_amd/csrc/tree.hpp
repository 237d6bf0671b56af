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
//     and a backward sweep, one worker per chunk) giving a reduced system on the
//     chunk ends with a direct end-to-end coupling; the reduced system is again a
//     forest of chains, so this repeats until every chain has at most two nodes;
//   * the final compressed forest is eliminated by height (a node pulls the
//     contributions of its children, so siblings never write the same parent) and
//     back-substituted by height;
//   * eliminated rows are recovered in parallel as x = alpha + beta*x_s + gamma*x_e
//     (one worker per chunk again, from the solution of the chunk's two ends).
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
  std::vector<int32_t> chunk_dyn;   // 1: the chunk's matrix values change between steps
  std::vector<int32_t> slot0;       // level-0 slot of each row (N entries)
  std::vector<int32_t> row0;        // row of each level-0 slot (-1 for padding)
  std::vector<int32_t> fin_par;     // final forest: parent (final-local) or -1
  std::vector<int32_t> fin_order;   // final-forest nodes ordered by height
  std::vector<int32_t> fin_hoff;    // height boundaries into fin_order, H+1 entries
  std::vector<int32_t> fin_cptr, fin_child;  // children of final-forest nodes (CSR)
  int heights = 0;                  // H
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
  int64_t total_nodes() const { return node_off.empty() ? 0 : node_off.back(); }
};

// LDS capacity for the staged top of the final forest (local entries).
constexpr int kTopLdsNodes = 1024;

// Builds the plan for a forest given by parent indices (par[i] < i or -1), with
// chains cut into chunks of `chunk` rows at level 0 and `chunk_hi` rows above.
// dyn_rows (optional, N entries): rows whose matrix values change between steps;
// chunks holding one, and the reduced chunks that depend on them, are refactored
// every step, the others reuse their factors. Within a level, chunks are placed
// (slot columns) dynamic first, then by decreasing length, so a wave's lanes take
// the same path and stay active together.
TreePlanHost build_tree_plan(const std::vector<int32_t>& par, int chunk, int chunk_hi,
                             const std::vector<uint8_t>* dyn_rows = nullptr);

// Splits the final forest of a plan for a launch of `groups` workgroups of
// `threads` threads (any number of workgroups can run it: workgroup b takes the
// groups b, b + nblocks, ...). The cut height minimises a latency model: one unit
// per pass of a workgroup over a height (ceil(nodes / threads)), up and down,
// plus two grid barriers when split; with `stage_top`, heights above a cut whose
// top fits in kTopLdsNodes LDS entries cost a tenth (the lists for that are
// built). No split when groups <= 1.
void partition_final_forest(TreePlanHost& P, int groups, int threads, bool stage_top = false);

// Parent of each row in an assembled CSR (the single column < row, or -1).
std::vector<int32_t> csr_parents(int N, const int32_t* rowptr, const int32_t* col);

struct TreePlanDev;
// Device view of a plan whose arrays are concatenated in one int32 buffer at
// `base`; with `flat` non-null the concatenation is appended there.
TreePlanDev plan_view(const TreePlanHost& P, const int32_t* base, std::vector<int32_t>* flat);

// Device view of a plan (pointers into one device allocation).
struct TreePlanDev {
  int levels, heights;
  int fin_cut, sub_groups;
  int top_n, top_ext;
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
  const int32_t* sub_order;
  const int32_t* sub_hoff;
  const int32_t* top_node;
  const int32_t* top_hoff;
  const int32_t* top_cptr;
  const int32_t* top_child;
  const int32_t* top_par;
};

// Per-value work arrays of one solve, over all levels (total_nodes each): d, up
// (row -> parent coupling), dn (parent -> row coupling), r, x; the elimination
// coefficients (three arrays a, b, g, used as (d', g', r') by the forward sweep
// and as (alpha, beta, gamma) afterwards); and the factors kept for the steps
// that reuse them: the forward multipliers f and the pivots' reciprocals.
template <class T>
struct TreeWork {
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
};

#if defined(__HIPCC__)
#define NRT_HD __host__ __device__ __forceinline__
#else
#define NRT_HD inline
#endif

// Rows of a chunk whose loads are issued together: the sweeps are recurrences, but
// their operands do not depend on them, so a batch of rows is fetched at once and
// the latency is paid once per batch rather than once per row.
constexpr int kTpsBatch = 4;

// Eliminate the interior of chunk `ch` of level `lv` and write its ends to lv+1.
template <class T>
NRT_HD void tps_reduce_chunk(const TreePlanDev& P, const TreeWork<T>& w, int lv, int ch) {
  T* __restrict__ d = w.d;
  T* __restrict__ up = w.up;
  T* __restrict__ dn = w.dn;
  T* __restrict__ r = w.r;
  T* __restrict__ ca = w.ca;
  T* __restrict__ cb = w.cb;
  T* __restrict__ cg = w.cg;
  T* __restrict__ cf = w.cf;
  T* __restrict__ ci = w.ci;
  const int NC = P.nchunk[lv];
  const int c = ch - P.chunk_off[lv];
  const int m = P.chunk_len[ch];
  const int base = P.node_off[lv] + c, nbase = P.node_off[lv + 1];
  if (m <= 2) {  // nothing to eliminate: the nodes pass through
    for (int k = 0; k < m; k++) {
      const int g = base + k * NC, t = nbase + P.next[g];
      d[t] = d[g];
      up[t] = up[g];
      dn[t] = dn[g];
      r[t] = r[g];
    }
    return;
  }
  // start row s = 0 and first interior row 1
  const T d0 = d[base], u0 = up[base], n0 = dn[base], r0 = r[base];
  const int ts = nbase + P.next[base];
  const int te = nbase + P.next[base + (m - 1) * NC];
  const int g1 = base + NC;
  const T dn1 = dn[g1];
  // forward sweep over rows k = 1..m-1:
  //   d'_k x_k + dn_{k+1} x_{k+1} + g_k x_s = r'_k,
  // stored as (ca, cb, cg) = (d', g, r') for interior rows; row m-1 becomes the
  // end row of the reduced system: G_e x_s + D_e x_e + (children) = R_e.
  T dp = d[g1], gk = up[g1], rp = r[g1];
  ca[g1] = dp;
  cb[g1] = gk;
  cg[g1] = rp;
  T pdp = dp, pgk = gk, prp = rp;  // values of the row before the last one
  T dnl = dn1;                       // dn of the last row
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
        br[j] = r[g];
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 + j;
      if (k <= m - 1) {
        pdp = dp;
        pgk = gk;
        prp = rp;
        const T f = bu[j] / dp;
        dp = bd[j] - f * bn[j];
        gk = -f * gk;
        rp = br[j] - f * rp;
        dnl = bn[j];
        const int g = base + k * NC;
        cf[g] = f;
        if (k <= m - 2) {
          ca[g] = dp;
          cb[g] = gk;
          cg[g] = rp;
        }
      }
    }
  }
  d[te] = dp;
  up[te] = gk;
  r[te] = rp;
  // backward sweep: x_k = alpha_k + beta_k x_s + gamma_k x_e, into (ca, cb, cg)
  T al, be, ga;
  {
    const int g2 = base + (m - 2) * NC;
    const T inv = T(1) / pdp;
    al = prp * inv;
    be = -pgk * inv;
    ga = -dnl * inv;
    ca[g2] = al;
    cb[g2] = be;
    cg[g2] = ga;
    ci[g2] = inv;
  }
  for (int k0 = m - 3; k0 >= 1; k0 -= kTpsBatch) {
    T bp[kTpsBatch], bg[kTpsBatch], brr[kTpsBatch], bn[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        const int g = base + k * NC;
        bp[j] = ca[g];
        bg[j] = cb[g];
        brr[j] = cg[g];
        bn[j] = dn[g + NC];  // coupling k -> k+1
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        const T inv = T(1) / bp[j];
        const T dcur = bn[j];
        al = (brr[j] - dcur * al) * inv;
        be = (-bg[j] - dcur * be) * inv;
        ga = -dcur * ga * inv;
        const int g = base + k * NC;
        ca[g] = al;
        cb[g] = be;
        cg[g] = ga;
        ci[g] = inv;
      }
    }
  }
  // start row s: up_s x_P + (d_s + dn_1 beta_1) x_s + dn_1 gamma_1 x_e = r_s - dn_1 alpha_1
  d[ts] = d0 + dn1 * be;
  up[ts] = u0;
  dn[ts] = n0;
  r[ts] = r0 - dn1 * al;
  dn[te] = dn1 * ga;  // s -> e coupling
}

// Right-hand side only, for a chunk whose factors (f, 1/pivot, beta, gamma, the
// reduced matrix values of its ends) are those of its last full elimination: the
// same operations on r as tps_reduce_chunk, in the same order, so the results are
// identical; the sweeps are multiply-adds with no division on the dependent path.
template <class T>
NRT_HD void tps_reduce_chunk_static(const TreePlanDev& P, const TreeWork<T>& w, int lv, int ch) {
  T* __restrict__ r = w.r;
  T* __restrict__ ca = w.ca;
  const T* __restrict__ dn = w.dn;
  const T* __restrict__ cf = w.cf;
  const T* __restrict__ ci = w.ci;
  const int NC = P.nchunk[lv];
  const int c = ch - P.chunk_off[lv];
  const int m = P.chunk_len[ch];
  const int base = P.node_off[lv] + c, nbase = P.node_off[lv + 1];
  if (m <= 2) {
    for (int k = 0; k < m; k++) {
      const int g = base + k * NC;
      r[nbase + P.next[g]] = r[g];
    }
    return;
  }
  const T r0 = r[base];
  const int ts = nbase + P.next[base];
  const int te = nbase + P.next[base + (m - 1) * NC];
  const T dn1 = dn[base + NC];
  // forward: rp_k = r_k - f_k rp_{k-1}
  T rp = r[base + NC];
  ca[base + NC] = rp;
  for (int k0 = 2; k0 <= m - 1; k0 += kTpsBatch) {
    T br[kTpsBatch], bf[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 + j;
      if (k <= m - 1) {
        const int g = base + k * NC;
        br[j] = r[g];
        bf[j] = cf[g];
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 + j;
      if (k <= m - 1) {
        rp = br[j] - bf[j] * rp;
        if (k <= m - 2) ca[base + k * NC] = rp;
      }
    }
  }
  r[te] = rp;
  // backward: alpha_k = (rp_k - dn_{k+1} alpha_{k+1}) / pivot_k
  const int g2 = base + (m - 2) * NC;
  T al = ca[g2] * ci[g2];
  ca[g2] = al;
  for (int k0 = m - 3; k0 >= 1; k0 -= kTpsBatch) {
    T bp[kTpsBatch], bi[kTpsBatch], bn[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        const int g = base + k * NC;
        bp[j] = ca[g];
        bi[j] = ci[g];
        bn[j] = dn[g + NC];
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 - j;
      if (k >= 1) {
        al = (bp[j] - bn[j] * al) * bi[j];
        ca[base + k * NC] = al;
      }
    }
  }
  r[ts] = r0 - dn1 * al;
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
    const T f = w.dn[c] / w.ca[c];
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
  w.x[base + v] = acc / w.ca[base + v];
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
  // carve from one buffer of top_lds_bytes<T>()
  NRT_HD static TopLds carve(void* p) {
    T* t = (T*)p;
    int32_t* i = (int32_t*)(t + 5 * kTopLdsNodes);
    return TopLds{t,     t + kTopLdsNodes,     t + 2 * kTopLdsNodes, t + 3 * kTopLdsNodes,
                  t + 4 * kTopLdsNodes, i, i + kTopLdsNodes + 1, i + 2 * kTopLdsNodes + 1,
                  i + 3 * kTopLdsNodes + 1};
  }
};
template <class T>
constexpr size_t top_lds_bytes() {
  return 5 * kTopLdsNodes * sizeof(T) + (4 * kTopLdsNodes + 2) * sizeof(int32_t);
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

template <class T>
NRT_HD void top_up(const TopLds<T>& L, int j) {
  T dv = L.d[j], rv = L.r[j];
  for (int k = L.cptr[j]; k < L.cptr[j + 1]; k++) {
    const int c = L.child[k];
    const T f = L.dn[c] / L.d[c];
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
  L.x[j] = acc / L.d[j];
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

template <int MODE, class T>
NRT_HD void tps_put(const TreePlanDev& P, const TreeWork<T>& w, int g, T v, double* out) {
  (void)P;
  if (MODE == OUT_LEVEL) w.x[g] = v;
  else out[g] = MODE == OUT_ADD ? out[g] + (double)v : (double)v;
}

// Back-substitution of chunk `ch` of level `lv` from the solution of level lv+1.
template <int MODE, class T>
NRT_HD void tps_back_chunk(const TreePlanDev& P, const TreeWork<T>& w, int lv, int ch,
                           double* out) {
  const int NC = P.nchunk[lv];
  const int c = ch - P.chunk_off[lv];
  const int m = P.chunk_len[ch];
  const int base = P.node_off[lv] + c, nbase = P.node_off[lv + 1];
  const T* __restrict__ xn = w.x + nbase;
  const int ge = base + (m - 1) * NC;
  const T xs = xn[P.next[base]];
  const T xe = m >= 2 ? xn[P.next[ge]] : xs;
  tps_put<MODE>(P, w, base, xs, out);
  if (m >= 2) tps_put<MODE>(P, w, ge, xe, out);
  const T* __restrict__ ca = w.ca;
  const T* __restrict__ cb = w.cb;
  const T* __restrict__ cg = w.cg;
  for (int k0 = 1; k0 <= m - 2; k0 += kTpsBatch) {
    T a[kTpsBatch], b[kTpsBatch], g[kTpsBatch];
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 + j;
      if (k <= m - 2) {
        const int gg = base + k * NC;
        a[j] = ca[gg];
        b[j] = cb[gg];
        g[j] = cg[gg];
      }
    }
#pragma unroll
    for (int j = 0; j < kTpsBatch; j++) {
      const int k = k0 + j;
      if (k <= m - 2) tps_put<MODE>(P, w, base + k * NC, a[j] + b[j] * xs + g[j] * xe, out);
    }
  }
}

// Level 0 as the final forest itself (a system with no chain longer than 2).
template <int MODE, class T>
NRT_HD void tps_final_out(const TreePlanDev& P, const TreeWork<T>& w, int v, double* out) {
  tps_put<MODE>(P, w, v, w.x[v], out);
}

}  // namespace nrt
