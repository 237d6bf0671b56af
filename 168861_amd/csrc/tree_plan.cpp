// Host-side plan of the tree partition solve (see tree.hpp).
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include "tree.hpp"
#include "../../include/neurite_hip.h"

namespace nrt {

std::vector<int32_t> csr_parents(int N, const int32_t* rowptr, const int32_t* col) {
  std::vector<int32_t> par(N, -1);
  for (int i = 0; i < N; i++) {
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      const int c = col[k];
      if (c < i) {
        if (par[i] >= 0 && par[i] != c)
          throw std::runtime_error("row " + std::to_string(i) + " couples to two earlier rows");
        par[i] = c;
      }
    }
  }
  return par;
}

// The top of the final forest for one workgroup (see TreePlanHost::top_*):
// the nodes of `is_top` (closed upwards) in height order, then their children
// outside the set. The lists are always built (the top can be solved from global
// memory); top_lds_ok says whether they fit max_entries LDS entries.
void stage_top_set(TreePlanHost& P, const std::vector<uint8_t>& is_top, int max_entries) {
  const int H = P.heights, n = (int)P.fin_par.size();
  std::vector<int32_t> loc(n, -1), node;
  int h0 = -1;
  P.top_hoff.clear();
  for (int k = 0; k < H; k++) {
    const int before = (int)node.size();
    for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++) {
      const int u = P.fin_order[q];
      if (!is_top[u]) continue;
      loc[u] = (int)node.size();
      node.push_back(u);
    }
    if (h0 < 0 && (int)node.size() > before) h0 = k;
    if (h0 >= 0) P.top_hoff.push_back(before);
  }
  const int nt = (int)node.size();
  if (h0 >= 0) P.top_hoff.push_back(nt);
  P.fin_cut = h0 < 0 ? H : h0;  // heights from here hold the top (and maybe local nodes)
  for (int j = 0; j < nt; j++) {
    const int v = node[j];
    for (int k = P.fin_cptr[v]; k < P.fin_cptr[v + 1]; k++) {
      const int c = P.fin_child[k];
      if (loc[c] < 0) {
        loc[c] = (int)node.size();
        node.push_back(c);
      }
    }
  }
  P.top_n = nt;
  P.top_ext = (int)node.size() - nt;
  P.top_node = node;
  P.top_cptr.assign(1, 0);
  P.top_child.clear();
  P.top_par.assign(nt, -1);
  for (int j = 0; j < nt; j++) {
    const int v = node[j];
    for (int k = P.fin_cptr[v]; k < P.fin_cptr[v + 1]; k++)
      P.top_child.push_back(loc[P.fin_child[k]]);
    P.top_cptr.push_back((int32_t)P.top_child.size());
    if (P.fin_par[v] >= 0) P.top_par[j] = loc[P.fin_par[v]];
  }
  P.top_lds_ok = nt > 0 && (int)node.size() <= max_entries ? 1 : 0;
  P.top_wide = 0;
  P.top_narrow = 0;
  P.top_kids = 0;
  for (int j = 0; j < nt; j++) P.top_kids = std::max(P.top_kids, P.top_cptr[j + 1] - P.top_cptr[j]);
  for (size_t k = 0; k + 1 < P.top_hoff.size(); k++) {
    const int wdt = P.top_hoff[k + 1] - P.top_hoff[k];
    P.top_wide = std::max(P.top_wide, wdt);
    if (wdt > 64) P.top_narrow = (int)k + 1;
  }
}

namespace {
struct LocalLevel {
  std::vector<int32_t> par;                 // level-local parents
  std::vector<uint8_t> dyn;                 // node values change between steps
  std::vector<int32_t> grp;                 // group of each node (local mode)
  std::vector<int32_t> cs, cl;              // chunks (local start, length)
  std::vector<uint8_t> cdyn;                // chunk holds a changing node
  std::vector<uint8_t> cwide;               // wide chunk (a whole chain above level 0)
  std::vector<int32_t> pos;                 // chunk -> slot column
  std::vector<int32_t> next;                // local next-level index
  std::vector<int32_t> slot;                // local node -> slot
  int nchunk = 0, cmax = 0, nslots = 0;
};
}  // namespace

// Chain tasks of the final forest (TreePlanHost::ftask_*). Children have larger
// indices than their parents, so one pass from the last node settles every task
// before its parent is seen.
static void build_final_tasks(TreePlanHost& P) {
  const int n = (int)P.fin_par.size();
  std::vector<int32_t> task(n, -1), th;  // task of each node; height of each task
  std::vector<std::vector<int32_t>> nodes;
  for (int u = n - 1; u >= 0; u--) {
    const int nc = P.fin_cptr[u + 1] - P.fin_cptr[u];
    if (nc == 1) {  // continues its only child's task (whose top that child is)
      const int c = P.fin_child[P.fin_cptr[u]];
      task[u] = task[c];
      nodes[task[u]].push_back(u);
    } else {
      int h = 0;
      for (int k = P.fin_cptr[u]; k < P.fin_cptr[u + 1]; k++)
        h = std::max(h, th[task[P.fin_child[k]]] + 1);
      task[u] = (int)nodes.size();
      nodes.push_back({u});
      th.push_back(h);
    }
  }
  const int nt = (int)nodes.size();
  int hn = 0;
  for (int t = 0; t < nt; t++) hn = std::max(hn, th[t] + 1);
  P.ftask_hn = n > 0 ? hn : 0;
  P.ftask_hoff.assign(P.ftask_hn + 1, 0);
  for (int t = 0; t < nt; t++) P.ftask_hoff[th[t] + 1]++;
  for (int h = 0; h < P.ftask_hn; h++) P.ftask_hoff[h + 1] += P.ftask_hoff[h];
  std::vector<int32_t> fill(P.ftask_hoff.begin(), P.ftask_hoff.end() - 1), order(nt);
  for (int t = 0; t < nt; t++) order[fill[th[t]]++] = t;
  P.ftask_off.assign(nt + 1, 0);
  P.ftask_node.clear();
  P.ftask_node.reserve(n);
  for (int i = 0; i < nt; i++) {
    const auto& v = nodes[order[i]];
    P.ftask_node.insert(P.ftask_node.end(), v.begin(), v.end());
    P.ftask_off[i + 1] = (int32_t)P.ftask_node.size();
  }
  // the same tasks in positions (the staged forest's layout)
  std::vector<int32_t> pos(n);
  for (int k = 0; k < n; k++) pos[P.ftask_node[k]] = k;
  P.ftask_par.assign(nt, -1);
  P.ftask_pnode.assign(nt, -1);
  P.ftask_cptr.assign(nt + 1, 0);
  P.ftask_child.clear();
  P.ftask_cnode.clear();
  for (int i = 0; i < nt; i++) {
    const int top = P.ftask_node[P.ftask_off[i + 1] - 1], bot = P.ftask_node[P.ftask_off[i]];
    if (P.fin_par[top] >= 0) {
      P.ftask_par[i] = pos[P.fin_par[top]];
      P.ftask_pnode[i] = P.fin_par[top];
    }
    for (int k = P.fin_cptr[bot]; k < P.fin_cptr[bot + 1]; k++) {
      P.ftask_child.push_back(pos[P.fin_child[k]]);
      P.ftask_cnode.push_back(P.fin_child[k]);
    }
    P.ftask_cptr[i + 1] = (int32_t)P.ftask_child.size();
  }
}

// Chain tasks of each group's local final nodes (TreePlanHost::ltask_*): as
// build_final_tasks, within a group (a local node's children are local nodes of
// its group), the groups' task lists side by side.
static void build_local_tasks(TreePlanHost& P, int G, int H) {
  const int n = (int)P.fin_par.size();
  std::vector<int32_t> task(n, -1);
  std::vector<std::vector<int32_t>> gth(G);                // task heights, per group
  std::vector<std::vector<std::vector<int32_t>>> gnodes(G);  // task nodes, per group
  int hn = 0;
  for (int g = 0; g < G; g++) {
    auto& th = gth[g];
    auto& nodes = gnodes[g];
    for (int k = P.loc_hoff[(size_t)g * H]; k < P.loc_hoff[(size_t)g * H + H]; k++) {
      const int u = P.loc_order[k];  // by height: children first
      const int nc = P.fin_cptr[u + 1] - P.fin_cptr[u];
      if (nc == 1) {
        const int c = P.fin_child[P.fin_cptr[u]];
        task[u] = task[c];
        nodes[task[u]].push_back(u);
      } else {
        int h = 0;
        for (int q = P.fin_cptr[u]; q < P.fin_cptr[u + 1]; q++)
          h = std::max(h, th[task[P.fin_child[q]]] + 1);
        task[u] = (int)nodes.size();
        nodes.push_back({u});
        th.push_back(h);
        hn = std::max(hn, h + 1);
      }
    }
  }
  P.ltask_hn = hn;
  P.ltask_hoff.assign((size_t)G * hn + 1, 0);
  P.ltask_off.assign(1, 0);
  P.ltask_node.clear();
  for (int g = 0; g < G; g++)
    for (int h = 0; h < hn; h++) {
      P.ltask_hoff[(size_t)g * hn + h] = (int32_t)P.ltask_off.size() - 1;
      for (size_t t = 0; t < gth[g].size(); t++) {
        if (gth[g][t] != h) continue;
        const auto& v = gnodes[g][t];
        P.ltask_node.insert(P.ltask_node.end(), v.begin(), v.end());
        P.ltask_off.push_back((int32_t)P.ltask_node.size());
      }
    }
  P.ltask_hoff[(size_t)G * hn] = (int32_t)P.ltask_off.size() - 1;
}

TreePlanHost build_tree_plan(const std::vector<int32_t>& par0, int chunk, int chunk_hi,
                             const std::vector<uint8_t>* dyn_rows,
                             const std::vector<int32_t>* groups, bool wide) {
  TreePlanHost P;
  P.C = std::max(3, chunk);
  P.C_hi = std::max(3, chunk_hi);
  const int N0 = (int)par0.size();
  for (int i = 0; i < N0; i++)
    if (par0[i] >= i) throw std::runtime_error("tree plan needs parent < child");
  std::vector<LocalLevel> lv;
  std::vector<int32_t> par = par0;
  std::vector<uint8_t> dyn(N0, 0);
  if (dyn_rows) {
    if ((int)dyn_rows->size() != N0) throw std::runtime_error("tree plan: dyn_rows size");
    dyn = *dyn_rows;
  }
  std::vector<int32_t> grp(N0, 0);
  int G = 0;
  if (groups) {
    if ((int)groups->size() != N0) throw std::runtime_error("tree plan: groups size");
    grp = *groups;
    for (int g : grp) {
      if (g < 0) throw std::runtime_error("tree plan: negative group");
      G = std::max(G, g + 1);
    }
  }
  P.groups = G;
  while (true) {
    LocalLevel L;
    L.par = par;
    L.dyn = dyn;
    L.grp = grp;
    const int n = (int)par.size();
    std::vector<int32_t> nchild(n, 0);
    for (int v = 0; v < n; v++)
      if (par[v] >= 0) nchild[par[v]]++;
    bool reduces = false;
    int v = 0;
    while (v < n) {
      int b = v;
      // a chain: consecutive single-child links, broken where the group changes
      while (b + 1 < n && par[b + 1] == b && nchild[b] == 1 && grp[b + 1] == grp[b]) b++;
      // above level 0 in local mode a long chain stays whole (wide chunks, which the
      // group-resident launch eliminates with the whole workgroup)
      int C = lv.empty() ? P.C : P.C_hi;
      const bool whole =
          wide && !lv.empty() && G > 0 && b - v + 1 > kWideMin && b - v + 1 <= kWideMax;
      if (whole) C = kWideMax;
      for (int a = v; a <= b; a += C) {
        const int m = std::min(C, b - a + 1);
        L.cs.push_back(a);
        L.cl.push_back(m);
        L.cwide.push_back(whole ? 1 : 0);
        if (m >= 3) reduces = true;
      }
      v = b + 1;
    }
    if (!reduces) {  // final forest
      L.cs.clear();
      L.cl.clear();
      lv.push_back(std::move(L));
      break;
    }
    L.next.assign(n, -1);
    int cnt = 0;
    for (size_t k = 0; k < L.cs.size(); k++) {
      const int a = L.cs[k], m = L.cl[k];
      L.next[a] = cnt++;
      if (m >= 2) L.next[a + m - 1] = cnt++;
    }
    // a chunk is dynamic when any of its nodes is; its surviving ends inherit that
    L.cdyn.assign(L.cs.size(), 0);
    std::vector<uint8_t> ndyn(cnt, 0);
    for (size_t k = 0; k < L.cs.size(); k++) {
      const int a = L.cs[k], m = L.cl[k];
      uint8_t f = 0;
      for (int j = a; j < a + m; j++) f |= L.dyn[j];
      L.cdyn[k] = f;
      ndyn[L.next[a]] = f;
      if (m >= 2) ndyn[L.next[a + m - 1]] = f;
    }
    dyn.swap(ndyn);
    std::vector<int32_t> ngrp(cnt, 0);
    for (size_t k = 0; k < L.cs.size(); k++) {
      const int a = L.cs[k], m = L.cl[k];
      ngrp[L.next[a]] = grp[a];
      if (m >= 2) ngrp[L.next[a + m - 1]] = grp[a];
    }
    grp.swap(ngrp);
    std::vector<int32_t> npar(cnt, -1);
    for (size_t k = 0; k < L.cs.size(); k++) {
      const int a = L.cs[k], m = L.cl[k];
      if (par[a] >= 0) {
        if (L.next[par[a]] < 0) throw std::runtime_error("tree plan: chunk parent eliminated");
        npar[L.next[a]] = L.next[par[a]];
      }
      if (m >= 2) npar[L.next[a + m - 1]] = L.next[a];
    }
    lv.push_back(std::move(L));
    par.swap(npar);
  }
  // slot maps
  const int nl = (int)lv.size();
  P.levels = nl - 1;
  for (int l = 0; l < nl; l++) {
    LocalLevel& L = lv[l];
    const int n = (int)L.par.size();
    L.slot.assign(n, -1);
    if (l == nl - 1) {
      for (int u = 0; u < n; u++) L.slot[u] = u;
      L.nslots = n;
    } else {
      L.nchunk = (int)L.cs.size();
      for (int m : L.cl) L.cmax = std::max(L.cmax, m);
      // slot columns: by group (local mode), then dynamic chunks first, then
      // longest first (stable)
      std::vector<int32_t> order(L.nchunk);
      for (int c = 0; c < L.nchunk; c++) order[c] = c;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int ga = L.grp[L.cs[a]], gb = L.grp[L.cs[b]];
        if (ga != gb) return ga < gb;
        if (L.cdyn[a] != L.cdyn[b]) return L.cdyn[a] > L.cdyn[b];
        return L.cl[a] > L.cl[b];
      });
      if (G > 0) {
        std::vector<int32_t> cnt(G + 1, 0);
        for (int c = 0; c < L.nchunk; c++) cnt[L.grp[L.cs[c]] + 1]++;
        for (int g = 0; g < G; g++) cnt[g + 1] += cnt[g];
        P.gcol.insert(P.gcol.end(), cnt.begin(), cnt.end());
      }
      L.pos.assign(L.nchunk, 0);
      for (int p = 0; p < L.nchunk; p++) L.pos[order[p]] = p;
      for (int c = 0; c < L.nchunk; c++)
        for (int k = 0; k < L.cl[c]; k++) L.slot[L.cs[c] + k] = k * L.nchunk + L.pos[c];
      L.nslots = L.cmax * L.nchunk;
    }
  }
  P.node_off.push_back(0);
  for (int l = 0; l < nl; l++) {
    P.nslots.push_back(lv[l].nslots);
    P.nchunk.push_back(lv[l].nchunk);
    P.node_off.push_back(P.node_off.back() + lv[l].nslots);
  }
  // flat slot-indexed arrays
  P.next.assign(P.total_nodes(), -1);
  P.chunk_off.push_back(0);
  for (int l = 0; l + 1 < nl; l++) {
    const LocalLevel& L = lv[l];
    const LocalLevel& Nx = lv[l + 1];
    const int base = P.node_off[l];
    for (size_t u = 0; u < L.par.size(); u++) {
      const int g = base + L.slot[u];
      if (L.next[u] >= 0) P.next[g] = Nx.slot[L.next[u]];
    }
    // chunk_len / chunk_dyn are indexed by slot column
    const size_t off = P.chunk_len.size();
    P.chunk_len.resize(off + L.nchunk);
    P.chunk_dyn.resize(off + L.nchunk);
    for (int c = 0; c < L.nchunk; c++) {
      P.chunk_len[off + L.pos[c]] = L.cl[c];
      // bit 0: eliminated in full every step; bit 1: wide (always in full)
      P.chunk_dyn[off + L.pos[c]] = L.cwide[c] ? 3 : L.cdyn[c];
    }
    P.chunk_off.push_back((int32_t)P.chunk_len.size());
  }
  P.slot0 = lv[0].slot;
  P.row0.assign(P.nslots[0], -1);
  for (int i = 0; i < N0; i++) P.row0[P.slot0[i]] = i;
  // level-0 rows whose right-hand side a step's update writes: those of the
  // dynamic chunks (a direct solve computes the others as it sweeps their chunks),
  // or every row when nothing is reduced
  P.dyn0.clear();
  if (P.levels == 0) {
    for (int v = 0; v < P.nslots[0]; v++)
      if (P.row0[v] >= 0) P.dyn0.push_back(v);
  } else {
    const int nc0 = P.chunk_off[1] - P.chunk_off[0];
    for (int c = 0; c < nc0; c++) {
      if (!(P.chunk_dyn[P.chunk_off[0] + c] & 1)) continue;
      for (int k = 0; k < P.chunk_len[P.chunk_off[0] + c]; k++) {
        const int v = k * P.nchunk[0] + c;
        if (P.row0[v] >= 0) P.dyn0.push_back(v);
      }
    }
  }
  // final forest (local index == slot): heights, children, order
  const std::vector<int32_t>& fp = lv[nl - 1].par;
  const int n = (int)fp.size();
  P.fin_par = fp;
  std::vector<int32_t> h(n, 0);
  for (int u = n - 1; u >= 0; u--)
    if (fp[u] >= 0) h[fp[u]] = std::max(h[fp[u]], h[u] + 1);
  int H = 0;
  for (int u = 0; u < n; u++) H = std::max(H, h[u] + 1);
  P.heights = H;
  P.fin_hoff.assign(H + 1, 0);
  for (int u = 0; u < n; u++) P.fin_hoff[h[u] + 1]++;
  for (int k = 0; k < H; k++) P.fin_hoff[k + 1] += P.fin_hoff[k];
  P.fin_order.assign(n, 0);
  std::vector<int32_t> fill(P.fin_hoff.begin(), P.fin_hoff.end() - 1);
  for (int u = 0; u < n; u++) P.fin_order[fill[h[u]]++] = u;
  P.fin_cptr.assign(n + 1, 0);
  for (int u = 0; u < n; u++)
    if (fp[u] >= 0) P.fin_cptr[fp[u] + 1]++;
  for (int u = 0; u < n; u++) P.fin_cptr[u + 1] += P.fin_cptr[u];
  P.fin_child.assign(P.fin_cptr[n], 0);
  std::vector<int32_t> pos(P.fin_cptr.begin(), P.fin_cptr.end() - 1);
  for (int u = 0; u < n; u++)
    if (fp[u] >= 0) P.fin_child[pos[fp[u]]++] = u;
  build_final_tasks(P);
  P.nchunk.push_back(0);  // keeps nchunk[levels] addressable
  if (G > 0) {
    // local final nodes: the whole subtree in the node's group (children have
    // larger indices, so one pass from the end settles every node)
    const std::vector<int32_t>& fg = lv[nl - 1].grp;
    std::vector<uint8_t> mixed(n, 0);
    for (int u = n - 1; u >= 0; u--)
      if (fp[u] >= 0 && (mixed[u] || fg[u] != fg[fp[u]])) mixed[fp[u]] = 1;
    P.loc_hoff.assign((size_t)G * H + 1, 0);
    for (int u = 0; u < n; u++)
      if (!mixed[u]) P.loc_hoff[(size_t)fg[u] * H + h[u] + 1]++;
    for (size_t i = 0; i + 1 < P.loc_hoff.size(); i++) P.loc_hoff[i + 1] += P.loc_hoff[i];
    P.loc_order.assign(P.loc_hoff.back(), 0);
    std::vector<int32_t> lf(P.loc_hoff.begin(), P.loc_hoff.end() - 1);
    for (int k = 0; k < H; k++)
      for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++) {
        const int u = P.fin_order[q];
        if (!mixed[u]) P.loc_order[lf[(size_t)fg[u] * H + h[u]]++] = u;
      }
    build_local_tasks(P, G, H);
    stage_top_set(P, mixed, kTopLdsLocal);
    // group-resident indices
    const int L = P.levels;
    P.lbase.assign((size_t)G * (L + 2), 0);
    std::vector<int32_t> lfin(n, -1);
    for (int g = 0; g < G; g++) {
      int32_t off = 0;
      for (int l = 1; l < L; l++) {
        const int32_t* gc = P.gcol.data() + (size_t)l * (G + 1);
        int cmax = 0;
        for (int c = gc[g]; c < gc[g + 1]; c++)
          cmax = std::max(cmax, P.chunk_len[P.chunk_off[l] + c]);
        P.lbase[(size_t)g * (L + 2) + l] = off;
        off += cmax * (gc[g + 1] - gc[g]);
      }
      P.lbase[(size_t)g * (L + 2) + L] = off;
      const int k0 = P.loc_hoff[(size_t)g * H], k1 = P.loc_hoff[(size_t)g * H + H];
      for (int k = k0; k < k1; k++) lfin[P.loc_order[k]] = off + (k - k0);
      off += k1 - k0;
      P.lbase[(size_t)g * (L + 2) + L + 1] = off;
      P.lds_nodes = std::max<int>(P.lds_nodes, off);
      int cols = 0;
      for (int l = 0; l < L; l++)
        cols += P.gcol[(size_t)l * (G + 1) + g + 1] - P.gcol[(size_t)l * (G + 1) + g];
      P.gl_cols = std::max(P.gl_cols, cols);
      P.gl_fin = std::max(P.gl_fin, k1 - k0);
    }
    auto final_code = [&](int u) { return mixed[u] ? -2 - u : lfin[u]; };
    P.lnext.assign(P.node_off[L], -1);
    for (int l = 0; l < L; l++) {
      const int NC = P.nchunk[l];
      const int32_t* gc = P.gcol.data() + (size_t)l * (G + 1);
      for (int g = 0; g < G; g++)
        for (int c = gc[g]; c < gc[g + 1]; c++) {
          const int m = P.chunk_len[P.chunk_off[l] + c];
          for (int e : {0, m - 1}) {
            const int slot = P.node_off[l] + c + e * NC, t = P.next[slot];
            int code;
            if (l + 1 < L) {
              const int NCn = P.nchunk[l + 1];
              const int32_t* gn = P.gcol.data() + (size_t)(l + 1) * (G + 1);
              code = P.lbase[(size_t)g * (L + 2) + l + 1] + (t / NCn) * (gn[g + 1] - gn[g]) +
                     (t % NCn - gn[g]);
            } else {
              code = final_code(t);
            }
            P.lnext[slot] = code;
          }
        }
    }
    const int nloc = (int)P.loc_order.size();
    P.lf_cptr.assign(nloc + 1, 0);
    P.lf_child.clear();
    P.lf_par.assign(nloc, -1);
    for (int k = 0; k < nloc; k++) {
      const int u = P.loc_order[k];
      for (int q = P.fin_cptr[u]; q < P.fin_cptr[u + 1]; q++)
        P.lf_child.push_back(lfin[P.fin_child[q]]);
      P.lf_cptr[k + 1] = (int32_t)P.lf_child.size();
      if (fp[u] >= 0) P.lf_par[k] = final_code(fp[u]);
    }
    for (int g = 0; g < G; g++) {
      const int k0 = P.loc_hoff[(size_t)g * H], k1 = P.loc_hoff[(size_t)g * H + H];
      P.gl_edges = std::max(P.gl_edges, P.lf_cptr[k1] - P.lf_cptr[k0]);
    }
    // wide chunks (levels above 0) of every group, by level
    P.wide_off.assign((size_t)G * (L + 1) + 1, 0);
    P.wide_col.clear();
    for (int g = 0; g < G; g++) {
      for (int l = 0; l < L; l++) {
        P.wide_off[(size_t)g * (L + 1) + l] = (int32_t)P.wide_col.size();
        if (l == 0) continue;
        const int32_t* gc = P.gcol.data() + (size_t)l * (G + 1);
        for (int c = gc[g]; c < gc[g + 1]; c++)
          if (P.chunk_dyn[P.chunk_off[l] + c] & 2) P.wide_col.push_back(c - gc[g]);
      }
      P.wide_off[(size_t)g * (L + 1) + L] = (int32_t)P.wide_col.size();
      P.gl_wide = std::max(P.gl_wide, P.wide_off[(size_t)g * (L + 1) + L] -
                                          P.wide_off[(size_t)g * (L + 1)]);
    }
    P.wide_off.back() = (int32_t)P.wide_col.size();
  }
  return P;
}

// Device view over one flat int32 buffer holding every plan array; `flat`
// receives the concatenation, the returned view has offsets relative to base.
// The plan's arrays in the order of the device buffer.
static void plan_parts(const TreePlanHost& P, const std::vector<int32_t>* (&parts)[kPlanParts]) {
  const std::vector<int32_t>* p[kPlanParts] = {
      &P.nslots,    &P.nchunk,    &P.node_off,  &P.next,      &P.chunk_off, &P.chunk_len,
      &P.chunk_dyn, &P.slot0,     &P.row0,      &P.fin_par,   &P.fin_order, &P.fin_hoff,
      &P.fin_cptr,  &P.fin_child, &P.ftask_hoff, &P.ftask_off, &P.ftask_node,
      &P.ftask_par, &P.ftask_cptr, &P.ftask_child, &P.ftask_pnode, &P.ftask_cnode, &P.dyn0,
      &P.ltask_hoff, &P.ltask_off, &P.ltask_node,
      &P.sub_order, &P.sub_hoff,  &P.top_node,  &P.top_hoff,
      &P.top_cptr,  &P.top_child, &P.top_par,   &P.gcol,      &P.loc_order, &P.loc_hoff,
      &P.lbase,     &P.lnext,     &P.lf_cptr,   &P.lf_child,  &P.lf_par,    &P.wide_off,
      &P.wide_col};
  for (int k = 0; k < kPlanParts; k++) parts[k] = p[k];
}

static TreePlanDev plan_from(const TreePlanHost& P, const int32_t* const* a) {
  const int ntask = P.ftask_off.empty() ? 0 : (int)P.ftask_off.size() - 1;
  return TreePlanDev{P.levels,     P.heights,    P.ftask_hn,  (int)P.dyn0.size(),
                     P.ltask_hn,   (int)P.fin_par.size(), ntask, (int)P.ftask_child.size(),
                     P.fin_cut,    P.sub_groups, P.top_n,     P.top_ext,
                     P.groups,     P.top_lds_ok, P.lds_nodes, P.top_wide,   P.top_narrow,
                     P.top_kids,   P.gl_cols,    P.gl_fin,    P.gl_edges,   P.gl_wide,
                     a[0],         a[1],         a[2],        a[3],         a[4],
                     a[5],         a[6],         a[7],        a[8],         a[9],
                     a[10],        a[11],        a[12],       a[13],        a[14],
                     a[15],        a[16],        a[17],       a[18],        a[19],
                     a[20],        a[21],        a[22],       a[23],        a[24],
                     a[25],        a[26],        a[27],       a[28],        a[29],
                     a[30],        a[31],        a[32],       a[33],        a[34],
                     a[35],        a[36],        a[37],       a[38],        a[39],
                     a[40],        a[41],        a[42]};
}

TreePlanDev plan_view(const TreePlanHost& P, const int32_t* base, std::vector<int32_t>* flat) {
  const std::vector<int32_t>* parts[kPlanParts];
  plan_parts(P, parts);
  const int32_t* a[kPlanParts];
  size_t pos = 0;
  for (int k = 0; k < kPlanParts; k++) {
    a[k] = base + pos;
    if (flat) {
      flat->insert(flat->end(), parts[k]->begin(), parts[k]->end());
      flat->push_back(0);
    }
    pos += parts[k]->size() + 1;
  }
  return plan_from(P, a);
}

TreePlanDev plan_view_parts(const TreePlanHost& P, std::vector<std::vector<int32_t>>& store) {
  const std::vector<int32_t>* parts[kPlanParts];
  plan_parts(P, parts);
  store.clear();
  store.reserve(kPlanParts);
  const int32_t* a[kPlanParts];
  for (int k = 0; k < kPlanParts; k++) {
    store.emplace_back(parts[k]->begin(), parts[k]->end());
    a[k] = store.back().data();
  }
  return plan_from(P, a);
}

SlotCsr slot_order_csr(const TreePlanHost& P, int N, const int32_t* rowptr, const int32_t* col,
                       const double* val, int H, const int32_t* hh_elem, const int32_t* hh_mat) {
  const int ns0 = P.nslots[0], nnz = rowptr[N];
  SlotCsr c;
  c.rowptr.assign(ns0 + 1, 0);
  c.col.assign(nnz, 0);
  c.val.assign(nnz, 0.0);
  c.role.assign(nnz, 0);
  c.hh_slot.assign(H, 0);
  c.hh_mat.assign(H, 0);
  for (int si = 0; si < ns0; si++) {
    const int row = P.row0[si];
    c.rowptr[si + 1] = c.rowptr[si] + (row >= 0 ? rowptr[row + 1] - rowptr[row] : 0);
  }
  for (int si = 0; si < ns0; si++) {
    const int row = P.row0[si];
    if (row < 0) continue;
    for (int k = rowptr[row], t = c.rowptr[si]; k < rowptr[row + 1]; k++, t++) {
      const int cc = col[k];
      c.col[t] = P.slot0[cc];
      c.val[t] = val[k];
      c.role[t] = cc == row ? 0 : (cc < row ? 1 : 2);
    }
  }
  for (int hh = 0; hh < H; hh++) {
    const int row = hh_elem[hh];
    c.hh_slot[hh] = P.slot0[row];
    c.hh_mat[hh] = c.rowptr[c.hh_slot[hh]] + (hh_mat[hh] - rowptr[row]);
  }
  return c;
}

// Top lists of the heights >= cut (see TreePlanHost::top_*), kept only when they
// fit in kTopLdsNodes LDS entries (a grid launch solves an unstaged top from the
// plan's global lists).
static void stage_final_top(TreePlanHost& P, int cut) {
  const int H = P.heights, n = (int)P.fin_par.size();
  std::vector<uint8_t> is_top(n, 0);
  for (int q = P.fin_hoff[cut]; q < P.fin_hoff[H]; q++) is_top[P.fin_order[q]] = 1;
  stage_top_set(P, is_top, kTopLdsNodes);
  P.fin_cut = cut;
  if (!P.top_lds_ok) {
    P.top_n = P.top_ext = 0;
    P.top_node.clear();
    P.top_hoff.clear();
    P.top_cptr.clear();
    P.top_child.clear();
    P.top_par.clear();
  }
}

void partition_final_forest(TreePlanHost& P, int groups, int threads, bool stage_top) {
  P.fin_cut = 0;
  P.sub_groups = 0;
  P.sub_order.clear();
  P.sub_hoff.clear();
  P.top_n = P.top_ext = 0;
  P.top_node.clear();
  P.top_hoff.clear();
  P.top_cptr.clear();
  P.top_child.clear();
  P.top_par.clear();
  const int H = P.heights, n = (int)P.fin_par.size();
  if (groups <= 1 || H <= 1 || n == 0) return;
  const int T = std::max(1, threads);
  std::vector<int32_t> h(n);
  for (int k = 0; k < H; k++)
    for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++) h[P.fin_order[q]] = k;
  // Cost units: one pass of a workgroup over a height through global memory; a
  // pass over LDS-staged heights costs kLdsPass of that, plus kStage to stage.
  constexpr int64_t kScale = 10, kLdsPass = 1, kStage = 2 * kScale, kBarriers = 4 * kScale;
  auto rounds = [&](int64_t nodes) { return (nodes + T - 1) / T; };
  auto top_cost = [&](int cut, int entries) {
    int64_t c = 0;
    const bool lds = stage_top && entries <= kTopLdsNodes;
    for (int k = cut; k < H; k++) {
      const int64_t r = rounds(P.fin_hoff[k + 1] - P.fin_hoff[k]);
      c += 2 * r * (lds ? kLdsPass : kScale);
    }
    return c + (lds ? kStage : 0);
  };
  int64_t best = top_cost(0, n);  // unsplit
  int best_cut = 0;
  std::vector<int32_t> root(n), best_root, best_bin;
  int best_G = 0;
  const int max_cut = std::min(H, 64);
  for (int cut = 1; cut <= max_cut; cut++) {
    // subtree root of every node below the cut: nodes by decreasing height, so a
    // parent's root is known before its children's
    std::vector<int32_t> roots;
    int ext = 0;
    for (int k = cut - 1; k >= 0; k--)
      for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++) {
        const int u = P.fin_order[q], p = P.fin_par[u];
        if (p < 0 || h[p] >= cut) {
          root[u] = u;
          roots.push_back(u);
          if (p >= 0) ext++;
        } else {
          root[u] = root[p];
        }
      }
    if ((int)roots.size() < 2) continue;
    // subtrees to groups, largest first onto the lightest group
    std::vector<int64_t> size(n, 0);
    for (int k = 0; k < cut; k++)
      for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++) size[root[P.fin_order[q]]]++;
    std::sort(roots.begin(), roots.end(), [&](int a, int b) {
      return size[a] != size[b] ? size[a] > size[b] : a < b;
    });
    const int G = std::min<int>(groups, (int)roots.size());
    std::vector<int64_t> load(G, 0);
    std::vector<int32_t> gof(n, -1);
    for (int u : roots) {
      int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      gof[u] = g;
      load[g] += size[u];
    }
    std::vector<int64_t> cnt((size_t)G * cut, 0);
    for (int k = 0; k < cut; k++)
      for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++)
        cnt[(size_t)gof[root[P.fin_order[q]]] * cut + k]++;
    int64_t cost = kBarriers;
    for (int k = 0; k < cut; k++) {
      int64_t mx = 0;
      for (int g = 0; g < G; g++) mx = std::max(mx, cnt[(size_t)g * cut + k]);
      cost += 2 * rounds(mx) * kScale;
    }
    cost += top_cost(cut, P.fin_hoff[H] - P.fin_hoff[cut] + ext);
    if (cost < best) {
      best = cost;
      best_cut = cut;
      best_root = root;
      best_bin.assign(n, -1);
      for (int u : roots) best_bin[u] = gof[u];
      best_G = G;
    }
  }
  if (best_cut > 0) {
    const int cut = best_cut, G = best_G;
    P.fin_cut = cut;
    P.sub_groups = G;
    P.sub_hoff.assign((size_t)G * cut + 1, 0);
    for (int k = 0; k < cut; k++)
      for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++) {
        const int u = P.fin_order[q];
        P.sub_hoff[(size_t)best_bin[best_root[u]] * cut + k + 1]++;
      }
    for (size_t i = 0; i + 1 < P.sub_hoff.size(); i++) P.sub_hoff[i + 1] += P.sub_hoff[i];
    P.sub_order.assign(P.fin_hoff[cut], 0);
    std::vector<int32_t> fill(P.sub_hoff.begin(), P.sub_hoff.end() - 1);
    for (int k = 0; k < cut; k++)
      for (int q = P.fin_hoff[k]; q < P.fin_hoff[k + 1]; q++) {
        const int u = P.fin_order[q];
        P.sub_order[fill[(size_t)best_bin[best_root[u]] * cut + k]++] = u;
      }
  }
  if (stage_top) stage_final_top(P, P.fin_cut);
}

std::vector<int32_t> row_groups(int N, int G) {
  std::vector<int32_t> g(N);
  for (int i = 0; i < N; i++) g[i] = (int32_t)((int64_t)i * G / N);
  return g;
}

void check_local_plan(const TreePlanHost& P) {
  auto need = [](bool ok, const char* what) {
    if (!ok) throw std::runtime_error(std::string("local plan: ") + what);
  };
  const int G = P.groups, H = P.heights, n = (int)P.fin_par.size(), L = P.levels;
  need(G >= 1 && L >= 1, "needs groups and a reduction level");
  need((int)P.gcol.size() == L * (G + 1), "gcol size");
  std::vector<int32_t> hgt(n);
  for (int h = 0; h < H; h++)
    for (int q = P.fin_hoff[h]; q < P.fin_hoff[h + 1]; q++) hgt[P.fin_order[q]] = h;
  // final nodes: each once, local (with its group and height) or top
  std::vector<int32_t> fgrp(n, -1), seen(n, 0);
  std::vector<uint8_t> top(n, 0);
  need((int)P.loc_hoff.size() == G * H + 1, "loc_hoff size");
  for (int g = 0; g < G; g++)
    for (int h = 0; h < H; h++)
      for (int k = P.loc_hoff[g * H + h]; k < P.loc_hoff[g * H + h + 1]; k++) {
        const int u = P.loc_order[k];
        need(hgt[u] == h, "local node listed at a wrong height");
        fgrp[u] = g;
        seen[u]++;
      }
  for (int j = 0; j < P.top_n; j++) {
    top[P.top_node[j]] = 1;
    seen[P.top_node[j]]++;
  }
  for (int u = 0; u < n; u++) need(seen[u] == 1, "final node placed twice or never");
  for (int u = 0; u < n; u++) {
    const int p = P.fin_par[u];
    if (top[u]) need(p < 0 || top[p], "top not closed upwards");
    else
      for (int k = P.fin_cptr[u]; k < P.fin_cptr[u + 1]; k++)
        need(!top[P.fin_child[k]] && fgrp[P.fin_child[k]] == fgrp[u],
             "a local node's child outside its group");
  }
  // group-resident indices: every chunk row and end, every local final node and
  // its links inside the group's range [0, lbase[g][L+1]), top codes on top nodes
  const bool gl = !P.lbase.empty();
  auto lcount = [&](int g) { return P.lbase[(size_t)g * (L + 2) + L + 1]; };
  if (gl) {
    need((int)P.lbase.size() == G * (L + 2) && (int)P.lnext.size() == P.node_off[L], "lbase size");
    for (int g = 0; g < G; g++) {
      need(lcount(g) <= P.lds_nodes, "group count above lds_nodes");
      for (int l = 1; l <= L; l++)
        need(P.lbase[(size_t)g * (L + 2) + l] <= P.lbase[(size_t)g * (L + 2) + l + 1],
             "lbase order");
      const int k0 = P.loc_hoff[(size_t)g * H], k1 = P.loc_hoff[(size_t)g * H + H];
      need(P.lbase[(size_t)g * (L + 2) + L] + (k1 - k0) == lcount(g), "local final block size");
      for (int k = k0; k < k1; k++) {
        for (int q = P.lf_cptr[k]; q < P.lf_cptr[k + 1]; q++)
          need(P.lf_child[q] >= P.lbase[(size_t)g * (L + 2) + L] && P.lf_child[q] < lcount(g),
               "local child index out of its group's final block");
        const int p = P.lf_par[k];
        need(p == -1 || (p >= 0 && p < lcount(g)) || (p <= -2 && -2 - p < n && top[-2 - p]),
             "local parent code");
      }
    }
  }
  // chunks: columns by group; both ends of a chunk go to its group (or the top)
  for (int lv = 0; lv < L; lv++) {
    const int NC = P.nchunk[lv], NCn = P.nchunk[lv + 1];
    const int32_t* gc = P.gcol.data() + (size_t)lv * (G + 1);
    need(gc[0] == 0 && gc[G] == NC, "gcol bounds");
    for (int g = 0; g < G; g++) {
      need(gc[g] <= gc[g + 1], "gcol order");
      for (int c = gc[g]; c < gc[g + 1]; c++) {
        const int m = P.chunk_len[P.chunk_off[lv] + c];
        for (int e : {0, m - 1}) {
          const int t = P.next[P.node_off[lv] + c + e * NC];
          need(t >= 0, "chunk end without a next slot");
          if (lv + 1 < L) {
            const int32_t* gn = P.gcol.data() + (size_t)(lv + 1) * (G + 1);
            const int cn = t % NCn;
            need(cn >= gn[g] && cn < gn[g + 1], "chunk end in another group's column");
          } else {
            need(top[t] || fgrp[t] == g, "chunk end in another group's final nodes");
          }
          if (gl) {
            const int code = P.lnext[P.node_off[lv] + c + e * NC];
            need((code >= 0 && code < lcount(g)) || (code <= -2 && -2 - code < n && top[-2 - code]),
                 "chunk end code");
          }
        }
        if (gl && lv >= 1) {
          const int base = P.lbase[(size_t)g * (L + 2) + lv] + (c - gc[g]);
          const int last = base + (m - 1) * (gc[g + 1] - gc[g]);
          need(base >= 0 && last < P.lbase[(size_t)g * (L + 2) + lv + 1], "chunk rows out of level");
        }
      }
    }
  }
}

}  // namespace nrt

using namespace nrt;

// Test hook (not part of the solver ABI): the tree partition solve of an assembled
// CSR system run on the host with the same per-chunk / per-node routines the
// device kernels execute, so CPU tests can check the plan logic against the
// oracle's sequential elimination. No solver handle ever calls it.
extern "C" int nrt_test_tree_solve_host(int N, const int32_t* rowptr, const int32_t* col,
                                        const double* val, const double* rhs, int chunk,
                                        int chunk_hi, const uint8_t* dyn_rows,
                                        const double* rhs_factor, int groups, double* x,
                                        int* levels, int* heights, int* cut) {
  try {
    std::vector<int32_t> par = csr_parents(N, rowptr, col);
    std::vector<uint8_t> dyn;
    if (dyn_rows) dyn.assign(dyn_rows, dyn_rows + N);
    // groups > 0: a grid launch's split of the final forest over that many
    // workgroups; groups < 0: local mode with -groups row groups, or with
    // -groups - 4096 groups in the group-resident form (a host buffer per group
    // stands in for its LDS)
    const bool local = groups < 0;
    const bool resident = groups <= -4096;
    std::vector<int32_t> rgrp;
    if (local) rgrp = row_groups(N, resident ? -groups - 4096 : -groups);
    TreePlanHost H = build_tree_plan(par, chunk, chunk_hi > 0 ? chunk_hi : chunk,
                                     dyn_rows ? &dyn : nullptr, local ? &rgrp : nullptr);
    if (local) check_local_plan(H);
    else partition_final_forest(H, groups, 256, groups > 0);
    if (!local) {
      // the split's promise to the device: a subtree node's children sit in its own
      // group at lower heights, and every node is placed exactly once
      const int n = (int)H.fin_par.size(), cut = H.fin_cut, G = H.sub_groups;
      std::vector<int32_t> grp(n, -1), hgt(n, -1), seen(n, 0);
      for (int g = 0; g < G; g++)
        for (int h = 0; h < cut; h++)
          for (int k = H.sub_hoff[g * cut + h]; k < H.sub_hoff[g * cut + h + 1]; k++) {
            grp[H.sub_order[k]] = g;
            hgt[H.sub_order[k]] = h;
            seen[H.sub_order[k]]++;
          }
      for (int h = cut; h < H.heights; h++)
        for (int k = H.fin_hoff[h]; k < H.fin_hoff[h + 1]; k++) seen[H.fin_order[k]]++;
      for (int u = 0; u < n; u++) {
        if (seen[u] != 1) throw std::runtime_error("final forest split: node placed twice or never");
        if (grp[u] < 0) continue;
        for (int k = H.fin_cptr[u]; k < H.fin_cptr[u + 1]; k++) {
          const int c = H.fin_child[k];
          if (grp[c] != grp[u] || hgt[c] >= hgt[u])
            throw std::runtime_error("final forest split: a child outside its parent's group");
        }
      }
    }
    {
      // the update's promise to a direct solve: every row is either in dyn0 (its
      // right-hand side written by the update) or in a static level-0 chunk (computed
      // by the sweep), never both
      std::vector<int> seen(H.nslots[0], 0);
      for (int v : H.dyn0) seen[v] += 1;
      if (H.levels > 0)
        for (int c = 0; c < H.chunk_off[1]; c++)
          if (!(H.chunk_dyn[c] & 1))
            for (int k = 0; k < H.chunk_len[c]; k++) seen[k * H.nchunk[0] + c] += 2;
      for (int v = 0; v < H.nslots[0]; v++)
        if ((H.row0[v] >= 0) != (seen[v] == 1 || seen[v] == 2))
          throw std::runtime_error("update rows: a row written twice or never");
    }
    {
      // the chain tasks' promise: every node once; above a task's bottom node, each
      // node's only child is the node below it; the bottom node's children sit in
      // tasks of lower task heights
      const int n = (int)H.fin_par.size();
      std::vector<int32_t> th(n, -1);
      for (int h = 0; h < H.ftask_hn; h++)
        for (int t = H.ftask_hoff[h]; t < H.ftask_hoff[h + 1]; t++)
          for (int k = H.ftask_off[t]; k < H.ftask_off[t + 1]; k++) {
            const int u = H.ftask_node[k];
            if (th[u] >= 0) throw std::runtime_error("chain tasks: node placed twice");
            th[u] = h;
            const int nc = H.fin_cptr[u + 1] - H.fin_cptr[u];
            if (k > H.ftask_off[t]) {
              if (nc != 1 || H.fin_child[H.fin_cptr[u]] != H.ftask_node[k - 1])
                throw std::runtime_error("chain tasks: a chain node without its one child below");
            } else {
              for (int q = H.fin_cptr[u]; q < H.fin_cptr[u + 1]; q++)
                if (th[H.fin_child[q]] < 0 || th[H.fin_child[q]] >= h)
                  throw std::runtime_error("chain tasks: a child not in a lower task height");
            }
          }
      for (int u = 0; u < n; u++)
        if (th[u] < 0) throw std::runtime_error("chain tasks: node never placed");
      // the task-indexed lists: per task, the parent of its top node and the children
      // of its bottom node, as nodes and as positions in ftask_node
      const int nt = H.ftask_hn > 0 ? H.ftask_hoff[H.ftask_hn] : 0;
      if ((int)H.ftask_par.size() != nt || (int)H.ftask_pnode.size() != nt ||
          (int)H.ftask_cptr.size() != nt + 1 || H.ftask_child.size() != H.ftask_cnode.size())
        throw std::runtime_error("task lists: sizes");
      std::vector<int32_t> pos(n, -1);
      for (int k = 0; k < n; k++) pos[H.ftask_node[k]] = k;
      for (int t = 0; t < nt; t++) {
        const int top = H.ftask_node[H.ftask_off[t + 1] - 1], bot = H.ftask_node[H.ftask_off[t]];
        const int p = H.fin_par[top];
        if (H.ftask_pnode[t] != p || H.ftask_par[t] != (p >= 0 ? pos[p] : -1))
          throw std::runtime_error("task lists: the parent of a task's top node");
        const int q0 = H.ftask_cptr[t], q1 = H.ftask_cptr[t + 1];
        if (q1 - q0 != H.fin_cptr[bot + 1] - H.fin_cptr[bot])
          throw std::runtime_error("task lists: the children of a task's bottom node");
        for (int q = q0; q < q1; q++) {
          const int c = H.fin_child[H.fin_cptr[bot] + (q - q0)];
          if (H.ftask_cnode[q] != c || H.ftask_child[q] != pos[c])
            throw std::runtime_error("task lists: the children of a task's bottom node");
        }
      }
    }
    if (local) {
      // the same for each group's local chain tasks: exactly its local final nodes,
      // a bottom node's children in its group at lower task heights
      const int n = (int)H.fin_par.size(), G = H.groups, TH = H.ltask_hn, Hh = H.heights;
      std::vector<int32_t> th(n, -1), grp(n, -1);
      for (int g = 0; g < G; g++)
        for (int k = H.loc_hoff[(size_t)g * Hh]; k < H.loc_hoff[(size_t)g * Hh + Hh]; k++)
          grp[H.loc_order[k]] = g;
      int placed = 0;
      for (int g = 0; g < G; g++)
        for (int h = 0; h < TH; h++)
          for (int t = H.ltask_hoff[g * TH + h]; t < H.ltask_hoff[g * TH + h + 1]; t++)
            for (int k = H.ltask_off[t]; k < H.ltask_off[t + 1]; k++) {
              const int u = H.ltask_node[k];
              if (grp[u] != g || th[u] >= 0)
                throw std::runtime_error("local chain tasks: a node placed twice or in another group");
              th[u] = h;
              placed++;
              const int nc = H.fin_cptr[u + 1] - H.fin_cptr[u];
              if (k > H.ltask_off[t]) {
                if (nc != 1 || H.fin_child[H.fin_cptr[u]] != H.ltask_node[k - 1])
                  throw std::runtime_error("local chain tasks: a chain node without its one child");
              } else {
                for (int q = H.fin_cptr[u]; q < H.fin_cptr[u + 1]; q++) {
                  const int c = H.fin_child[q];
                  if (grp[c] != g || th[c] < 0 || th[c] >= h)
                    throw std::runtime_error("local chain tasks: a child not below in its group");
                }
              }
            }
      if (placed != (int)H.loc_order.size())
        throw std::runtime_error("local chain tasks: a local node never placed");
    }
    // every plan array and every work array in an allocation of its own, of exactly
    // its length, so that a bounds-checking build of this hook catches an access
    // past any one of them (the device keeps them in one buffer each)
    std::vector<std::vector<int32_t>> parts;
    TreePlanDev P = plan_view_parts(H, parts);
    const size_t T = (size_t)H.total_nodes();
    std::vector<std::vector<double>> wa(11, std::vector<double>(T, 0.0));
    TreeWork<double> w{wa[0].data(), wa[1].data(), wa[2].data(), wa[3].data(),
                       wa[4].data(), wa[5].data(), wa[6].data(), wa[7].data(),
                       wa[8].data(), wa[9].data(), wa[10].data()};
    const int32_t* s0 = H.slot0.data();
    for (int i = 0; i < N; i++) {
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int c = col[k];
        if (c == i) w.d[s0[i]] += val[k];
        else if (c < i) w.up[s0[i]] = val[k];
        else w.dn[s0[c]] = val[k];
      }
    }
    const bool reuse = dyn_rows != nullptr;
    if (reuse) {
      // full elimination with another right-hand side: factors for the static chunks
      for (int i = 0; i < N; i++) w.r[s0[i]] = rhs_factor ? rhs_factor[i] : 0.0;
      for (int lv = 0; lv < H.levels; lv++)
        for (int ch = H.chunk_off[lv]; ch < H.chunk_off[lv + 1]; ch++)
          tps_reduce_chunk(P, w, lv, ch);
    }
    // with reuse, two solves in a row (as consecutive steps do): the second must
    // not depend on what the first left behind
    std::vector<double> xs(H.nslots[0], 0.0);  // the solution in level-0 slot order
    // local mode: the top of the final forest (LDS-staged when it fits, with host
    // memory standing in for LDS)
    auto solve_top = [&]() {
      if (H.top_n > 0 && H.top_lds_ok) {
        std::vector<unsigned char> lds(top_lds_bytes<double>(gl_top_entries(P)));
        const TopLds<double> L = TopLds<double>::carve(lds.data(), gl_top_entries(P));
        for (int j = 0; j < top_stage_count(P); j++) top_stage(P, w, L, j);
        const int nh = H.heights - H.fin_cut;
        for (int h = 0; h < nh; h++)
          for (int j = L.hoff[h]; j < L.hoff[h + 1]; j++) top_up(L, j);
        for (int h = nh - 1; h >= 0; h--)
          for (int j = L.hoff[h]; j < L.hoff[h + 1]; j++) top_down(L, j);
        for (int j = 0; j < H.top_n; j++) top_store(P, w, L, j);
      } else if (H.top_n > 0) {
        const int nh = H.heights - H.fin_cut;
        for (int h = 0; h < nh; h++)
          for (int j = H.top_hoff[h]; j < H.top_hoff[h + 1]; j++) tps_final_up(P, w, H.top_node[j]);
        for (int h = nh - 1; h >= 0; h--)
          for (int j = H.top_hoff[h]; j < H.top_hoff[h + 1]; j++)
            tps_final_down(P, w, H.top_node[j]);
      }
    };
    // group-resident form: every group's values and staged metadata in a buffer
    // of its own, laid out as the launch lays out its LDS (one top-input buffer)
    const GlIO<double> io{w, w, w, false};
    std::vector<std::vector<double>> lbuf;
    const GlLayout lay = gl_layout<double>(P, false);
    auto gl_view = [&](int g, TreeWorkL<double>& lw, GlMeta& M) {
      unsigned char* b = (unsigned char*)lbuf[g].data();
      lw = gl_carve<double>(b + lay.top, H.lds_nodes);
      M = gl_meta_carve(b + lay.top + lay.values, P);
    };
    // wide chunks: every PCR step reads all rows, then writes all rows (as the
    // launch's workgroup does between its barriers)
    auto wide_reduce = [&](const TreeWorkL<double>& lw, const GlMeta& M, int lv, int j) {
      const int q = 4 * (M.lvl[7 * lv + 3] + j);
      const int m = M.col[q], es = M.col[q + 1], ee = M.col[q + 2];
      const int base = M.lvl[7 * lv + 2] + j, st = M.lvl[7 * lv + 1];
      for (int k = 1; k <= m - 2; k++) wide_init(lw, base, st, m, k);
      std::vector<WideRow<double>> tmp(m);
      for (int s = 0, h = 1; s < wide_steps(m - 2); s++, h *= 2) {
        for (int k = 1; k <= m - 2; k++) tmp[k] = wide_step(lw, base, st, m, k, h);
        for (int k = 1; k <= m - 2; k++) wide_put(lw, base, st, k, tmp[k]);
      }
      for (int k = 1; k <= m - 2; k++) wide_finish(lw, base, st, k);
      wide_ends(lw, base, st, m, es, ee, io.ends(lw, M.scal[0]));
    };
    auto wide_back = [&](const TreeWorkL<double>& lw, const GlMeta& M, int lv, int j) {
      const int q = 4 * (M.lvl[7 * lv + 3] + j);
      const int m = M.col[q], es = M.col[q + 1], ee = M.col[q + 2];
      const int base = M.lvl[7 * lv + 2] + j, st = M.lvl[7 * lv + 1];
      const EndsLocal<double> E = io.ends(lw, M.scal[0]);
      wide_back_rows(lw, base, st, m, E.x(es), E.x(ee), 0, 1);
    };
    auto gl_levels = [&](int g, int how) {
      TreeWorkL<double> lw;
      GlMeta M;
      gl_view(g, lw, M);
      for (int j = 0; j < M.lvl[1]; j++) glm_reduce<true>(io, lw, M, 0, j, how);
      for (int lv = 1; lv < H.levels; lv++) {
        for (int j = 0; j < M.lvl[7 * lv + 1]; j++) glm_reduce<false>(io, lw, M, lv, j, how);
        for (int i = 0; i < M.lvl[7 * lv + 6]; i++)
          wide_reduce(lw, M, lv, M.wcol[M.lvl[7 * lv + 5] + i]);
      }
    };
    for (int pass = 0; pass < (reuse ? 2 : 1); pass++) {
      for (int i = 0; i < N; i++) w.r[s0[i]] = rhs[i];
      for (int lv = 0; lv < H.levels; lv++)
        for (int ch = H.chunk_off[lv]; ch < H.chunk_off[lv + 1]; ch++) {
          if (reuse && !H.chunk_dyn[ch]) tps_reduce_chunk_static(P, w, lv, ch);
          else tps_reduce_chunk(P, w, lv, ch);
        }
      if (resident) {
        const int G = H.groups, L = H.levels, Hh = H.heights;
        if (lbuf.empty()) {
          lbuf.assign(G, std::vector<double>(lay.total / sizeof(double) + 1));
          for (int g = 0; g < G; g++) {
            TreeWorkL<double> lw;
            GlMeta M;
            gl_view(g, lw, M);
            gl_meta_stage(P, g, M, 0, 1);
          }
          // factors first (with another right-hand side when reusing them)
          for (int i = 0; i < N; i++) w.r[s0[i]] = rhs_factor ? rhs_factor[i] : 0.0;
          for (int g = 0; g < G; g++) gl_levels(g, kGlFull);
          for (int i = 0; i < N; i++) w.r[s0[i]] = rhs[i];
        }
        for (int g = G - 1; g >= 0; g--) {
          gl_levels(g, reuse ? kGlStep : kGlFull);
          TreeWorkL<double> lw;
          GlMeta M;
          gl_view(g, lw, M);
          for (int h = 0; h < Hh; h++)
            for (int j = M.hoff[h]; j < M.hoff[h + 1]; j++) glm_final_up(io, lw, M, j);
        }
        solve_top();
        for (int g = 0; g < G; g++) {
          TreeWorkL<double> lw;
          GlMeta M;
          gl_view(g, lw, M);
          for (int h = Hh - 1; h >= 0; h--)
            for (int j = M.hoff[h]; j < M.hoff[h + 1]; j++) glm_final_down(io, lw, M, j);
          for (int lv = L - 1; lv >= 1; lv--) {
            for (int j = 0; j < M.lvl[7 * lv + 1]; j++)
              glm_back<false, OUT_LEVEL>(io, lw, M, lv, j, nullptr);
            for (int i = 0; i < M.lvl[7 * lv + 6]; i++)
              wide_back(lw, M, lv, M.wcol[M.lvl[7 * lv + 5] + i]);
          }
          for (int j = 0; j < M.lvl[1]; j++) glm_back<true, OUT_ASSIGN>(io, lw, M, 0, j, xs.data());
        }
        continue;
      }
      if (local) {
        // local mode in the order of its launch: every group alone up to its local
        // final nodes (here last group first: no group may need another's work),
        // the top, then every group back down
        const int G = H.groups;
        for (int g = G - 1; g >= 0; g--) {
          for (int lv = 0; lv < H.levels; lv++) {
            const int* gc = H.gcol.data() + (size_t)lv * (G + 1);
            for (int ch = H.chunk_off[lv] + gc[g]; ch < H.chunk_off[lv] + gc[g + 1]; ch++) {
              if (reuse && !H.chunk_dyn[ch]) tps_reduce_chunk_static(P, w, lv, ch);
              else tps_reduce_chunk(P, w, lv, ch);
            }
          }
          const int TH = H.ltask_hn;
          for (int h = 0; h < TH; h++)
            for (int t = H.ltask_hoff[g * TH + h]; t < H.ltask_hoff[g * TH + h + 1]; t++)
              tps_task_up(P, w, P.ltask_off, P.ltask_node, t);
        }
        solve_top();
        for (int g = 0; g < G; g++) {
          const int TH = H.ltask_hn;
          for (int h = TH - 1; h >= 0; h--)
            for (int t = H.ltask_hoff[g * TH + h]; t < H.ltask_hoff[g * TH + h + 1]; t++)
              tps_task_down(P, w, P.ltask_off, P.ltask_node, t);
          for (int lv = H.levels - 1; lv >= 0; lv--) {
            const int* gc = H.gcol.data() + (size_t)lv * (G + 1);
            for (int ch = H.chunk_off[lv] + gc[g]; ch < H.chunk_off[lv] + gc[g + 1]; ch++) {
              if (lv > 0) tps_back_chunk<OUT_LEVEL>(P, w, lv, ch, nullptr);
              else tps_back_chunk<OUT_ASSIGN>(P, w, 0, ch, xs.data());
            }
          }
        }
        continue;
      }
      // final forest in the order of a grid launch: the groups' subtrees below the
      // cut, then the heights above it, then back down
      const int cut = H.fin_cut, G = H.sub_groups;
      for (int h = 0; h < cut; h++)
        for (int g = 0; g < G; g++)
          for (int k = H.sub_hoff[g * cut + h]; k < H.sub_hoff[g * cut + h + 1]; k++)
            tps_final_up(P, w, H.sub_order[k]);
      if (H.top_n > 0) {
        // the LDS-staged top, with host memory standing in for LDS
        std::vector<unsigned char> lds(top_lds_bytes<double>());
        const TopLds<double> L = TopLds<double>::carve(lds.data());
        for (int j = 0; j < top_stage_count(P); j++) top_stage(P, w, L, j);
        const int nh = H.heights - cut;
        for (int h = 0; h < nh; h++)
          for (int j = L.hoff[h]; j < L.hoff[h + 1]; j++) top_up(L, j);
        for (int h = nh - 1; h >= 0; h--)
          for (int j = L.hoff[h]; j < L.hoff[h + 1]; j++) top_down(L, j);
        for (int j = 0; j < H.top_n; j++) top_store(P, w, L, j);
      } else if (cut == 0 && H.ftask_hn > 0) {
        // unsplit: the forest staged at task positions (host memory of exactly its
        // size standing in for LDS), then by chain tasks from the work arrays; the
        // two must give the same solution to the bit
        std::vector<unsigned char> lds(forest_lds_bytes<double>(P));
        const ForestLds<double> L = ForestLds<double>::carve(lds.data(), P);
        for (int j = 0; j < forest_stage_count(P); j++) forest_stage(P, w, L, j);
        for (int h = 0; h < H.ftask_hn; h++)
          for (int t = L.thoff[h]; t < L.thoff[h + 1]; t++) forest_up(L, t);
        for (int h = H.ftask_hn - 1; h >= 0; h--)
          for (int t = L.thoff[h]; t < L.thoff[h + 1]; t++) forest_down(L, t);
        for (int j = 0; j < P.fin_n; j++) forest_store(P, w, L, j);
        const int fb = H.node_off[H.levels];
        const std::vector<double> staged(w.x + fb, w.x + fb + P.fin_n);
        for (int h = 0; h < H.ftask_hn; h++)
          for (int t = H.ftask_hoff[h]; t < H.ftask_hoff[h + 1]; t++) tps_ftask_up(P, w, t);
        for (int h = H.ftask_hn - 1; h >= 0; h--)
          for (int t = H.ftask_hoff[h]; t < H.ftask_hoff[h + 1]; t++) tps_ftask_down(P, w, t);
        if (std::memcmp(staged.data(), w.x + fb, sizeof(double) * P.fin_n) != 0)
          throw std::runtime_error("staged final forest: a solution differs from the chain tasks'");
      } else {
        for (int h = cut; h < H.heights; h++)
          for (int k = H.fin_hoff[h]; k < H.fin_hoff[h + 1]; k++)
            tps_final_up(P, w, H.fin_order[k]);
        for (int h = H.heights - 1; h >= cut; h--)
          for (int k = H.fin_hoff[h]; k < H.fin_hoff[h + 1]; k++)
            tps_final_down(P, w, H.fin_order[k]);
      }
      for (int h = cut - 1; h >= 0; h--)
        for (int g = 0; g < G; g++)
          for (int k = H.sub_hoff[g * cut + h]; k < H.sub_hoff[g * cut + h + 1]; k++)
            tps_final_down(P, w, H.sub_order[k]);
      for (int lv = H.levels - 1; lv >= 1; lv--)
        for (int ch = H.chunk_off[lv]; ch < H.chunk_off[lv + 1]; ch++)
          tps_back_chunk<OUT_LEVEL>(P, w, lv, ch, xs.data());
      if (H.levels == 0) {
        for (int v = 0; v < H.nslots[0]; v++) tps_final_out<OUT_ASSIGN>(P, w, v, xs.data());
      } else {
        for (int ch = H.chunk_off[0]; ch < H.chunk_off[1]; ch++)
          tps_back_chunk<OUT_ASSIGN>(P, w, 0, ch, xs.data());
      }
    }
    for (int i = 0; i < N; i++) x[i] = xs[s0[i]];
    if (levels) *levels = H.levels;
    if (heights) *heights = H.heights;
    if (cut)
      *cut = resident ? H.lds_nodes
           : local  ? H.top_n | (H.top_lds_ok ? (1 << 30) : 0)
                   : H.fin_cut | (H.top_n > 0 ? (1 << 30) : 0);
    return NRT_OK;
  } catch (const std::exception&) {
    return NRT_ERR_TOPOLOGY;
  }
}
