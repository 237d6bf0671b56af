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
//   * eliminated rows are recovered in parallel as x = alpha + beta*x_s + gamma*x_e.
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
  int C = 32;                       // chunk length
  int levels = 0;                   // reduction levels (L)
  std::vector<int32_t> nslots;      // slots per level, L+1 entries (last = final forest)
  std::vector<int32_t> nchunk;      // chunks per reduction level (L entries, +1 padding)
  std::vector<int32_t> node_off;    // offset of each level in the flat arrays, L+2
  std::vector<int32_t> next;        // flat: slot in level+1 of a surviving node (-1 otherwise)
  std::vector<int32_t> es, ee;      // flat: level+1 slots of the chunk ends of an eliminated node
  std::vector<int32_t> chunk_off;   // per level, L+1 entries, into chunk_len
  std::vector<int32_t> chunk_len;
  std::vector<int32_t> slot0;       // level-0 slot of each row (N entries)
  std::vector<int32_t> fin_par;     // final forest: parent (final-local) or -1
  std::vector<int32_t> fin_order;   // final-forest nodes ordered by height
  std::vector<int32_t> fin_hoff;    // height boundaries into fin_order, H+1 entries
  std::vector<int32_t> fin_cptr, fin_child;  // children of final-forest nodes (CSR)
  int heights = 0;                  // H
  int64_t total_nodes() const { return node_off.empty() ? 0 : node_off.back(); }
};

// Builds the plan for a forest given by parent indices (par[i] < i or -1).
TreePlanHost build_tree_plan(const std::vector<int32_t>& par, int chunk);

// Parent of each row in an assembled CSR (the single column < row, or -1).
std::vector<int32_t> csr_parents(int N, const int32_t* rowptr, const int32_t* col);

struct TreePlanDev;
// Device view of a plan whose arrays are concatenated in one int32 buffer at
// `base`; with `flat` non-null the concatenation is appended there.
TreePlanDev plan_view(const TreePlanHost& P, const int32_t* base, std::vector<int32_t>* flat);

// Device view of a plan (pointers into one device allocation).
struct TreePlanDev {
  int levels, heights;
  const int32_t* nslots;
  const int32_t* nchunk;
  const int32_t* node_off;
  const int32_t* next;
  const int32_t* es;
  const int32_t* ee;
  const int32_t* chunk_off;
  const int32_t* chunk_len;
  const int32_t* slot0;
  const int32_t* fin_par;
  const int32_t* fin_order;
  const int32_t* fin_hoff;
  const int32_t* fin_cptr;
  const int32_t* fin_child;
};

// Per-value work arrays of one solve, over all levels (total_nodes each): d, up
// (row -> parent coupling), dn (parent -> row coupling), r, x; and the
// elimination coefficients (three arrays a, b, g, used as (d', g', r') by the
// forward sweep and as (alpha, beta, gamma) afterwards).
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
};

#if defined(__HIPCC__)
#define NRT_HD __host__ __device__ __forceinline__
#else
#define NRT_HD inline
#endif

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
  // forward sweep over interior rows k = 1..m-2: d'_k x_k + dn_{k+1} x_{k+1} + g_k x_0 = r'_k
  int g = base + NC;
  T dp = d[g], gk = up[g], rp = r[g];
  ca[g] = dp;
  cb[g] = gk;
  cg[g] = rp;
  // values of the next row are loaded one step ahead
  T nu = up[g + NC], nd = d[g + NC], ndn = dn[g + NC], nr = r[g + NC];
  for (int k = 2; k <= m - 2; k++) {
    g += NC;
    const T u0 = nu, d0 = nd, dn0 = ndn, r0 = nr;
    nu = up[g + NC];
    nd = d[g + NC];
    ndn = dn[g + NC];
    nr = r[g + NC];
    const T f = u0 / dp;
    dp = d0 - f * dn0;
    gk = -f * gk;
    rp = r0 - f * rp;
    ca[g] = dp;
    cb[g] = gk;
    cg[g] = rp;
  }
  // end row e = m-1 eliminates x_{m-2}: G_e x_s + D_e x_e + children = R_e
  const int ge = base + (m - 1) * NC;
  const int te = nbase + P.next[ge];
  {
    const T f = nu / dp;  // nu, nd, ndn, nr hold row m-1 now
    d[te] = nd - f * ndn;
    up[te] = -f * gk;
    r[te] = nr - f * rp;
  }
  // backward sweep: x_k = alpha_k + beta_k x_s + gamma_k x_e
  T al, be, ga;
  {
    const int g2 = base + (m - 2) * NC;
    const T inv = T(1) / ca[g2];
    al = cg[g2] * inv;
    be = -cb[g2] * inv;
    ga = -ndn * inv;
    ca[g2] = al;
    cb[g2] = be;
    cg[g2] = ga;
  }
  T dnx = dn[base + (m - 2) * NC];
  for (int k = m - 3; k >= 1; k--) {
    const int g2 = base + k * NC;
    const T inv = T(1) / ca[g2];
    const T r2 = cg[g2], gg = cb[g2];
    const T dcur = dnx;
    dnx = dn[g2];
    al = (r2 - dcur * al) * inv;
    be = (-gg - dcur * be) * inv;
    ga = -dcur * ga * inv;
    ca[g2] = al;
    cb[g2] = be;
    cg[g2] = ga;
  }
  // start row s: up_s x_P + (d_s + dn_1 beta_1) x_s + dn_1 gamma_1 x_e = r_s - dn_1 alpha_1
  {
    const T dn1 = dn[base + NC];
    const int ts = nbase + P.next[base];
    d[ts] = d[base] + dn1 * be;
    up[ts] = up[base];
    dn[ts] = dn[base];
    r[ts] = r[base] - dn1 * al;
    dn[te] = dn1 * ga;  // s -> e coupling
  }
}

// Final forest, elimination of node v (all of its children are final already).
template <class T>
NRT_HD void tps_final_up(const TreePlanDev& P, const TreeWork<T>& w, int v) {
  const int base = P.node_off[P.levels];
  T dv = w.d[base + v], rv = w.r[base + v];
  for (int k = P.fin_cptr[v]; k < P.fin_cptr[v + 1]; k++) {
    const int c = base + P.fin_child[k];
    const T f = w.dn[c] / w.d[c];
    dv = dv - f * w.up[c];
    rv = rv - f * w.r[c];
  }
  w.d[base + v] = dv;
  w.r[base + v] = rv;
}

template <class T>
NRT_HD void tps_final_down(const TreePlanDev& P, const TreeWork<T>& w, int v) {
  const int base = P.node_off[P.levels];
  const int p = P.fin_par[v];
  T acc = w.r[base + v];
  if (p >= 0) acc = acc - w.up[base + v] * w.x[base + p];
  w.x[base + v] = acc / w.d[base + v];
}

// Back-substitution of slot v of level lv from level lv+1.
template <class T>
NRT_HD void tps_back(const TreePlanDev& P, const TreeWork<T>& w, int lv, int v) {
  const int base = P.node_off[lv], nbase = P.node_off[lv + 1];
  const int g = base + v;
  const int nx = P.next[g];
  if (nx >= 0) {
    w.x[g] = w.x[nbase + nx];
  } else {
    const int s = P.es[g];
    if (s < 0) return;  // padding slot
    const T xs = w.x[nbase + s];
    const T xe = w.x[nbase + P.ee[g]];
    w.x[g] = w.ca[g] + w.cb[g] * xs + w.cg[g] * xe;
  }
}

}  // namespace nrt
