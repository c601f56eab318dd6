// Host-side set-up of the FVT step (see fvt_prep.h).
//
// Every coefficient is formed with the same operations, in the same order,
// as the reference set-up routine cited beside it, so that the device step
// reproduces the reference arithmetic.  Topology searches that the reference
// does by linear scans (is_edge.f90, in_array*.f90) use hash maps here; their
// results are identical.
#include "fvt_prep.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <queue>
#include <unordered_map>

namespace storm {

namespace {

constexpr double HALF = 0.5;
constexpr double ONE = 1.0;
constexpr double ZERO = 0.0;
const double ONE_THIRD = 1.0 / 3.0;

inline uint64_t pair_key(int a, int b) {
  uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
  return ((uint64_t)lo << 32) | hi;
}

// cde_t2t.f90:30-47 (inner sweep trip count fixed at sweep start).
void cde_t2t(int* row, int& n) {
  int l = 1;
  for (;;) {
    int j = row[l];
    const int mend = n;
    for (int m = l + 1; m <= mend; ++m) {
      int k = row[m];
      if (j == k) {
        n = n - 1;
        for (int nn = m; nn <= n; ++nn) row[nn] = row[nn + 1];
      }
    }
    l = l + 1;
    if (l > n) break;
  }
}

// elemnt_db.f90:42-128 on the edge table (ee: e(1), e(2) per edge).
int molecule(int N, int E, const std::vector<int>& ee, std::vector<int>& ptr,
             std::vector<int>& cells, std::string& err) {
  const int MAXD = 20, W = MAXD + 2;
  if (N < 4) { err = "too few triangles in computational domain (< 4)"; return STORM_E_MESH; }
  std::vector<int> w((size_t)N * W, 0);
  auto row = [&](int i) { return &w[(size_t)i * W]; };
  for (int i = 0; i < E; ++i) {
    int e1 = ee[2 * i], e2 = ee[2 * i + 1];
    if (e1 < 0 || e2 < 0) continue;
    int* r1 = row(e1); int* r2 = row(e2);
    if (r1[0] >= 3 || r2[0] >= 3) { err = "triangle connectivity table incorrect"; return STORM_E_MESH; }
    r1[++r1[0]] = e2;
    r2[++r2[0]] = e1;
  }
  for (int i = 0; i < N; ++i) {
    int* ri = row(i);
    const int n = ri[0];
    if (n == 3) continue;
    if (n < 1 || n > 3) { err = "triangle connectivity table incorrect"; return STORM_E_MESH; }
    for (int l = 1; l <= n; ++l) {
      int j = ri[l];
      const int* rj = row(j);
      const int mm = std::min(3, rj[0]);
      for (int m = 1; m <= mm; ++m) {
        int k = rj[m];
        if (k == i) continue;
        if (ri[0] + 1 > MAXD) { err = "molecule overflow"; return STORM_E_MESH; }
        ri[++ri[0]] = k;
      }
    }
    int c = ri[0];
    cde_t2t(ri, c);
    ri[0] = c;
  }
  for (int i = 0; i < N; ++i) {
    int* ri = row(i);
    const int n = ri[0];
    if (n >= 3) continue;
    int j = ri[n];
    const int* rj = row(j);
    const int mm = rj[0];
    for (int m = 1; m <= mm; ++m) {
      int k = rj[m];
      if (k == i) continue;
      if (ri[0] + 1 > MAXD) { err = "molecule overflow"; return STORM_E_MESH; }
      ri[++ri[0]] = k;
    }
    int c = ri[0];
    cde_t2t(ri, c);
    ri[0] = c;
  }
  ptr.assign(N + 1, 0);
  for (int i = 0; i < N; ++i) ptr[i + 1] = ptr[i] + row(i)[0];
  cells.resize(ptr[N]);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < row(i)[0]; ++k) cells[ptr[i] + k] = row(i)[k + 1];
  return STORM_OK;
}

// elemnt_db.f90:136-214: the higher-order least-squares molecule t2tHD
// (ACTIVATE HDLS).  Row i: the other triangles on each of i's vertices, in
// n2t order, then cde_t2t.  Rows left with fewer than three are augmented as
// :168-197 write it: the loop index j = 2..nn names a TRIANGLE (grid(j)) and
// a candidate is skipped when its 1-based column in the n2t row equals j.
// Triangle numbers count from the first cell of i's reach (base), so that a
// reach in a batch gets the molecule it gets alone.
int hd_molecule(int N, const int32_t* tri, const std::vector<int>& n2t_ptr,
                const std::vector<int>& n2t, const std::vector<int>& base,
                std::vector<int>& ptr, std::vector<int>& cells, std::string& err) {
  int maxc = 0;
  for (size_t p = 0; p + 1 < n2t_ptr.size(); ++p) maxc = std::max(maxc, n2t_ptr[p + 1] - n2t_ptr[p]);
  const int W = 4 * maxc + 3;  // columns, the count included
  std::vector<int> w((size_t)N * W, 0);
  auto row = [&](int i) { return &w[(size_t)i * W]; };
  auto push = [&](int* r, int t) -> bool {
    if (r[0] + 2 > W) return false;
    r[++r[0]] = t;
    return true;
  };
  for (int i = 0; i < N; ++i) {
    int* ri = row(i);
    for (int j = 0; j < 3; ++j) {
      const int k = tri[3 * i + j];
      for (int t = n2t_ptr[k]; t < n2t_ptr[k + 1]; ++t)
        if (n2t[t] != i && !push(ri, n2t[t])) { err = "allocation bounds in T2TW"; return STORM_E_MESH; }
    }
    int c = ri[0];
    cde_t2t(ri, c);
    ri[0] = c;
  }
  for (int i = 0; i < N; ++i) {
    int* ri = row(i);
    const int n = ri[0];
    if (n > 2) continue;
    if (n < 1) { err = "triangle connectivity table incorrect"; return STORM_E_MESH; }
    const int nn = std::min(3, n + 1);
    for (int j = 2; j <= nn; ++j) {
      const int tj = base[i] + j - 1;
      if (tj >= N) { err = "triangle connectivity table incorrect"; return STORM_E_MESH; }
      for (int k = 0; k < 3; ++k) {
        const int l = tri[3 * tj + k];
        const int cl = n2t_ptr[l + 1] - n2t_ptr[l];
        for (int m = 2; m <= cl + 1; ++m) {
          const int t = n2t[n2t_ptr[l] + m - 2];
          if (t == i || m == j) continue;
          if (!push(ri, t)) { err = "allocation bounds in T2TW"; return STORM_E_MESH; }
        }
      }
    }
    int c = ri[0];
    cde_t2t(ri, c);
    ri[0] = c;
  }
  ptr.assign(N + 1, 0);
  for (int i = 0; i < N; ++i) ptr[i + 1] = ptr[i] + row(i)[0];
  cells.resize(ptr[N]);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < row(i)[0]; ++k) cells[ptr[i] + k] = row(i)[k + 1];
  return STORM_OK;
}

// Edge table: given, with the swap rule of tecplot_read.f90:139-142, or built.
int edge_table(const storm_mesh_t* m, std::vector<int>& ep, std::vector<int>& ee,
               std::string& err) {
  if (m->n_edges > 0) {
    if (!m->edge_p || !m->edge_e) { err = "n_edges > 0 but edge_p/edge_e missing"; return STORM_E_ARG; }
    ep.assign(m->edge_p, m->edge_p + 2 * (size_t)m->n_edges);
    ee.assign(m->edge_e, m->edge_e + 2 * (size_t)m->n_edges);
    for (int k = 0; k < m->n_edges; ++k)
      if (ee[2 * k] < 0) { ee[2 * k] = ee[2 * k + 1]; ee[2 * k + 1] = -1; }
    for (int k = 0; k < m->n_edges; ++k) {
      if (ep[2 * k] < 0 || ep[2 * k] >= m->n_pts || ep[2 * k + 1] < 0 || ep[2 * k + 1] >= m->n_pts ||
          ee[2 * k] < 0 || ee[2 * k] >= m->n_elems || ee[2 * k + 1] >= m->n_elems) {
        err = "edge table entry out of range"; return STORM_E_MESH;
      }
    }
  } else {
    build_edges(m->n_elems, m->n_pts, m->tri, ep, ee);
  }
  return STORM_OK;
}

// Cell order along a Hilbert curve over the centroids (one scale for both
// axes, 2^16 cells a side): runs of consecutive cells are compact patches,
// which is what the stage kernel's tiles are.
uint64_t hilbert_index(uint32_t x, uint32_t y) {
  const uint32_t n = 1u << 16;
  uint64_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) { x = n - 1 - x; y = n - 1 - y; }
      std::swap(x, y);
    }
  }
  return d;
}

// Reverse Cuthill-McKee order of cells [a, a + n) over their edge neighbours
// (t2t3): from a pseudo-peripheral cell (the last one reached by a
// breadth-first search from the first, repeated while the depth grows),
// breadth-first with each cell's unvisited neighbours by ascending degree,
// then reversed; every connected part in turn.  Returns local ids.
std::vector<int> rcm_order(int a, int n, const std::vector<int>& t2t3) {
  auto nbrs = [&](int i, int* out) {
    int c = 0;
    for (int j = 0; j < 3; ++j) {
      const int k = t2t3[3 * (size_t)(a + i) + j];
      if (k >= a && k < a + n) out[c++] = k - a;
    }
    return c;
  };
  std::vector<int> deg(n);
  int tmp[3];
  for (int i = 0; i < n; ++i) deg[i] = nbrs(i, tmp);
  std::vector<int> order, lvl(n, -1);
  order.reserve(n);
  std::vector<char> seen(n, 0);
  auto bfs_last = [&](int s, int& depth) {  // the deepest cell of s's part, least degree
    std::vector<int> q{s};
    std::fill(lvl.begin(), lvl.end(), -1);
    lvl[s] = 0;
    int last = s;
    for (size_t h = 0; h < q.size(); ++h) {
      const int i = q[h];
      const int c = nbrs(i, tmp);
      for (int j = 0; j < c; ++j)
        if (lvl[tmp[j]] < 0) { lvl[tmp[j]] = lvl[i] + 1; q.push_back(tmp[j]); }
      if (lvl[i] > lvl[last] || (lvl[i] == lvl[last] && deg[i] < deg[last])) last = i;
    }
    depth = lvl[last];
    return last;
  };
  for (int s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    int depth = -1, d2 = 0, s = s0;
    for (int it = 0; it < 8; ++it) {
      const int t = bfs_last(s, d2);
      if (d2 <= depth) break;
      depth = d2;
      s = t;
    }
    const size_t b = order.size();
    order.push_back(s);
    seen[s] = 1;
    for (size_t h = b; h < order.size(); ++h) {
      const int c = nbrs(order[h], tmp);
      std::sort(tmp, tmp + c, [&](int x, int y) { return deg[x] != deg[y] ? deg[x] < deg[y] : x < y; });
      for (int j = 0; j < c; ++j)
        if (!seen[tmp[j]]) { seen[tmp[j]] = 1; order.push_back(tmp[j]); }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

std::vector<int> hilbert_order(int N, const std::vector<double>& xc, const std::vector<double>& yc) {
  double x0 = xc[0], x1 = xc[0], y0 = yc[0], y1 = yc[0];
  for (int i = 1; i < N; ++i) {
    x0 = std::min(x0, xc[i]); x1 = std::max(x1, xc[i]);
    y0 = std::min(y0, yc[i]); y1 = std::max(y1, yc[i]);
  }
  double span = std::max(x1 - x0, y1 - y0);
  if (!(span > 0.0)) span = 1.0;
  const double f = 65535.0 / span;
  std::vector<std::pair<uint64_t, int>> key(N);
  for (int i = 0; i < N; ++i) {
    const uint32_t xi = (uint32_t)std::min(65535.0, std::max(0.0, (xc[i] - x0) * f));
    const uint32_t yi = (uint32_t)std::min(65535.0, std::max(0.0, (yc[i] - y0) * f));
    key[i] = {hilbert_index(xi, yi), i};
  }
  std::sort(key.begin(), key.end());
  std::vector<int> order(N);
  for (int i = 0; i < N; ++i) order[i] = key[i].second;
  return order;
}

}  // namespace

void build_edges(int N, int P, const int32_t* tri, std::vector<int>& ep, std::vector<int>& ee) {
  (void)P;
  std::unordered_map<uint64_t, int> idx;
  idx.reserve((size_t)N * 2);
  ep.clear();
  ee.clear();
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) {
      int a = tri[3 * i + j], b = tri[3 * i + (j + 1) % 3];
      auto it = idx.find(pair_key(a, b));
      if (it == idx.end()) {
        idx.emplace(pair_key(a, b), (int)(ep.size() / 2));
        ep.push_back(a); ep.push_back(b);
        ee.push_back(i); ee.push_back(-1);
      } else {
        ee[2 * it->second + 1] = i;
      }
    }
}

int build_molecule(const storm_mesh_t* m, std::vector<int>& ptr, std::vector<int>& cells,
                   std::string& err) {
  std::vector<int> ep, ee;
  int rc = edge_table(m, ep, ee, err);
  if (rc) return rc;
  return molecule(m->n_elems, (int)(ep.size() / 2), ee, ptr, cells, err);
}

namespace {

// solverFVT.f90:244-255 sweeps the cells in input order; cell i takes the
// new values of its t2t neighbours j < i and the old values of those j > i.
// A cell's level is one above every cell it must follow: each j < i of its
// row (read new) and each j < i whose row names i (it reads i's old value,
// so it runs first).  The cells of one level read and write disjoint values,
// so a level-by-level sweep gives the serial sweep's result bit for bit.
// Rows never cross reaches, so every reach's levels start at 0.
void rsavg_schedule(Prep& S, const std::vector<int>& tp, const std::vector<int>& tt,
                    const std::vector<int>& cell_ptr) {
  const int N = S.N, NR = (int)cell_ptr.size() - 1;
  std::vector<int> lev(N, 0);
  for (int i = 0; i < N; ++i) {
    int l = lev[i];
    for (int t = tp[i]; t < tp[i + 1]; ++t)
      if (tt[t] < i) l = std::max(l, lev[tt[t]] + 1);
    lev[i] = l;
    for (int t = tp[i]; t < tp[i + 1]; ++t)
      if (tt[t] > i) lev[tt[t]] = std::max(lev[tt[t]], l + 1);
  }
  S.r_lev.assign(NR + 1, 0);
  for (int r = 0; r < NR; ++r) {
    int top = 0;
    for (int i = cell_ptr[r]; i < cell_ptr[r + 1]; ++i) top = std::max(top, lev[i] + 1);
    S.r_lev[r + 1] = S.r_lev[r] + top;
  }
  std::vector<int> reach_of(N);
  for (int r = 0; r < NR; ++r)
    for (int i = cell_ptr[r]; i < cell_ptr[r + 1]; ++i) reach_of[i] = r;
  const int L = S.r_lev[NR];
  S.rs_lptr.assign(L + 1, 0);
  for (int i = 0; i < N; ++i) ++S.rs_lptr[S.r_lev[reach_of[i]] + lev[i] + 1];
  for (int l = 0; l < L; ++l) S.rs_lptr[l + 1] += S.rs_lptr[l];
  std::vector<int> fill(S.rs_lptr.begin(), S.rs_lptr.end() - 1);
  S.rs_cell.assign(N, 0);
  // within a level: rows of at most four entries first (the device sweep
  // gives longer rows a second pass that whole waves of short rows skip),
  // each group by ascending internal id
  for (int pass = 0; pass < 2; ++pass)
    for (int ii = 0; ii < N; ++ii) {
      const int i = S.perm[ii];
      if ((tp[i + 1] - tp[i] > 4) != (pass == 1)) continue;
      S.rs_cell[fill[S.r_lev[reach_of[i]] + lev[i]]++] = ii;
    }
  S.rs_nptr.assign(N + 1, 0);
  S.rs_nbr.clear();
  for (int k = 0; k < N; ++k) {
    const int i = S.perm[S.rs_cell[k]];
    for (int t = tp[i]; t < tp[i + 1]; ++t) {
      const int j = tt[t];
      S.rs_nbr.push_back(S.iperm[j] | (j > i ? (int)0x80000000u : 0));
    }
    S.rs_nptr[k + 1] = (int)S.rs_nbr.size();
  }
}

// The tables of the shared-memory sweep (see Prep::rs_fast); false when the
// mesh exceeds its limits (the sweep then reads global memory per level).
bool rsavg_fast_tables(Prep& S) {
  const int N = S.N, NR = (int)S.r_lev.size() - 1, L = S.r_lev[NR];
  const int W = RSF_W, G = RSF_G;
  S.rs_fast = false;
  S.rs_code.clear(); S.rs_farw.clear(); S.rs_fptr.clear(); S.rs_fsrc.clear();
  S.rs_fdst.clear(); S.rs_pptr.clear(); S.rs_psrc.clear(); S.rs_pdst.clear(); S.rs_hdr.clear();
  for (int l = 0; l < L; ++l)
    if (S.rs_lptr[l + 1] - S.rs_lptr[l] > RSF_NT) return false;
  for (int k = 0; k < N; ++k)
    if (S.rs_nptr[k + 1] - S.rs_nptr[k] > RSF_E) return false;
  std::vector<int> pos(N), levp(N), lreach(L);
  for (int k = 0; k < N; ++k) pos[S.rs_cell[k]] = k;
  for (int r = 0; r < NR; ++r)
    for (int l = S.r_lev[r]; l < S.r_lev[r + 1]; ++l) {
      lreach[l] = r;
      for (int k = S.rs_lptr[l]; k < S.rs_lptr[l + 1]; ++k) levp[k] = l;
    }
  auto base = [&](int k) { return S.rs_lptr[S.r_lev[lreach[levp[k]]]]; };
  auto rend = [&](int k) { return S.rs_lptr[S.r_lev[lreach[levp[k]] + 1]]; };
  // the level whose fill overwrites position p's ring slot (p + W's fill)
  const int NEVER = std::numeric_limits<int>::max();
  auto evict = [&](int p) {
    const int q = p + W;
    return q < rend(p) ? levp[q] - G - 1 : NEVER;
  };
  S.rs_code.assign((size_t)RSF_E * N, -1);
  S.rs_farw.assign(N, -1);
  // far values: new ones (written by their producer), old ones (filled from
  // the previous sweep before their first reader)
  std::vector<int> fn_end(N, -1), fo_first(N, NEVER), fo_last(N, -1);
  std::vector<std::pair<int, int>> far_entries;  // (code index, position), kind by sign
  for (int k = 0; k < N; ++k) {
    const int lc = levp[k];
    for (int m = S.rs_nptr[k], e = 0; m < S.rs_nptr[k + 1]; ++m, ++e) {
      const int code = S.rs_nbr[m];
      const bool old = code < 0;
      const int p = pos[code & 0x7fffffff];
      const int lj = levp[p];
      const bool ring = (!old || lj <= lc + G) && evict(p) > lc;
      const size_t ci = (size_t)e * N + k;
      if (ring) {
        S.rs_code[ci] = (p - base(p)) % W;
      } else if (!old) {
        fn_end[p] = std::max(fn_end[p], lc);
        far_entries.push_back({(int)ci, p});
      } else {
        fo_first[p] = std::min(fo_first[p], lc);
        fo_last[p] = std::max(fo_last[p], lc);
        far_entries.push_back({(int)ci, -1 - p});
      }
    }
  }
  // old far values: filled at the latest level before their first reader
  // with room in its fill list (else by the reach's prologue, level -1)
  std::vector<int> fcount(L, 0);  // (the ring fills take their own lanes)
  std::vector<int> fo_level(N, -2);
  for (int p = 0; p < N; ++p) {
    if (fo_last[p] < 0) continue;
    const int l0 = S.r_lev[lreach[levp[p]]];
    int lf = fo_first[p] - 1;
    while (lf >= l0 && fcount[lf] >= RSF_NT) --lf;
    if (lf >= l0) { ++fcount[lf]; fo_level[p] = lf; }
    else fo_level[p] = -1;
  }
  // far slots: intervals [start, end] of levels, greedy in start order; a
  // slot is free again for an interval starting after its last reader
  struct Iv { int start, end, p; bool old; };
  std::vector<Iv> iv;
  for (int p = 0; p < N; ++p) {
    if (fn_end[p] >= 0) iv.push_back({levp[p], fn_end[p], p, false});
    if (fo_last[p] >= 0) {
      const int l0 = S.r_lev[lreach[levp[p]]];
      iv.push_back({fo_level[p] >= 0 ? fo_level[p] : l0 - 1, fo_last[p], p, true});
    }
  }
  std::sort(iv.begin(), iv.end(), [](const Iv& a, const Iv& b) {
    return a.start != b.start ? a.start < b.start : a.p < b.p;
  });
  std::vector<int> fn_slot(N, -1), fo_slot(N, -1);
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>,
                      std::greater<std::pair<int, int>>> busy;  // (end, slot)
  std::vector<int> free_slots;
  int nslots = 0;
  for (const Iv& v : iv) {
    while (!busy.empty() && busy.top().first < v.start) {
      free_slots.push_back(busy.top().second);
      busy.pop();
    }
    int slot;
    if (!free_slots.empty()) { slot = free_slots.back(); free_slots.pop_back(); }
    else slot = nslots++;
    if (nslots > RSF_FS) return false;
    busy.push({v.end, slot});
    (v.old ? fo_slot : fn_slot)[v.p] = slot;
  }
  for (const auto& fe : far_entries) {
    const int p = fe.second >= 0 ? fe.second : -1 - fe.second;
    S.rs_code[fe.first] = W + (fe.second >= 0 ? fn_slot[p] : fo_slot[p]);
  }
  for (int p = 0; p < N; ++p)
    if (fn_slot[p] >= 0) S.rs_farw[p] = fn_slot[p];
  // per level, the old far values it fills; per reach, its prologue: the
  // ring slots of its first G + 1 levels and the old far values of level -1
  std::vector<std::vector<int>> fl(L);
  std::vector<std::vector<int>> pro(NR);
  for (int p = 0; p < N; ++p) {
    if (fo_last[p] < 0) continue;
    if (fo_level[p] >= 0) fl[fo_level[p]].push_back(p);
    else pro[lreach[levp[p]]].push_back(p);
  }
  S.rs_fptr.assign(L + 1, 0);
  for (int l = 0; l < L; ++l) {
    for (int p : fl[l]) { S.rs_fsrc.push_back(p); S.rs_fdst.push_back(fo_slot[p]); }
    S.rs_fptr[l + 1] = (int)S.rs_fsrc.size();
  }
  S.rs_pptr.assign(NR + 1, 0);
  for (int r = 0; r < NR; ++r) {
    const int l0 = S.r_lev[r], l1 = S.r_lev[r + 1];
    for (int l = l0; l < std::min(l1, l0 + G + 1); ++l)
      for (int k = S.rs_lptr[l]; k < S.rs_lptr[l + 1]; ++k) {
        S.rs_psrc.push_back(k);
        S.rs_pdst.push_back((k - S.rs_lptr[l0]) % W);
      }
    for (int p : pro[r]) { S.rs_psrc.push_back(p); S.rs_pdst.push_back(W + fo_slot[p]); }
    S.rs_pptr[r + 1] = (int)S.rs_psrc.size();
  }
  S.rs_hdr.assign(8 * ((size_t)L + 1), 0);
  for (int l = 0; l < L; ++l) {
    int* h = &S.rs_hdr[8 * (size_t)l];
    h[0] = S.rs_lptr[l];
    h[1] = S.rs_lptr[l + 1] - S.rs_lptr[l];
    const int lf = l + G + 1;
    if (lf < S.r_lev[lreach[l] + 1]) { h[2] = S.rs_lptr[lf]; h[3] = S.rs_lptr[lf + 1] - S.rs_lptr[lf]; }
    h[4] = S.rs_fptr[l];
    h[5] = S.rs_fptr[l + 1] - S.rs_fptr[l];
  }
  S.rs_nfar = nslots;
  S.rs_fast = true;
  return true;
}

}  // namespace

int wind_coefs(const Prep& S, int64_t n, const double* c, const double* wx, const double* wy,
               std::vector<double>& w, std::string& err) {
  const int N = S.N;
  if (n != S.N && n != S.P) {
    err = "wind data must hold one value per cell or one per vertex";
    return STORM_E_ARG;
  }
  if (!c || !wx || !wy) { err = "null wind array"; return STORM_E_ARG; }
  constexpr double RHO = 998.60;  // constants.f90:29
  w.assign(2 * (size_t)N, ZERO);
  for (int ii = 0; ii < N; ++ii) {
    const int i = S.perm[ii];
    double w0, w1, w2;
    if (n == S.N) {
      w0 = c[i]; w1 = wx[i]; w2 = wy[i];
    } else {  // vtx2ctr.f90:32-35 on the triangle's vertices in input order
      const int a = S.vperm[S.cvtx[ii]], b = S.vperm[S.cvtx[N + ii]],
                d = S.vperm[S.cvtx[2 * (size_t)N + ii]];
      w0 = (c[a] + c[b] + c[d]) * ONE_THIRD;
      w1 = (wx[a] + wx[b] + wx[d]) * ONE_THIRD;
      w2 = (wy[a] + wy[b] + wy[d]) * ONE_THIRD;
    }
    const double mag = std::sqrt(w1 * w1 + w2 * w2);
    w[2 * (size_t)ii] = w0 * w1 * mag / RHO;
    w[2 * (size_t)ii + 1] = w0 * w2 * mag / RHO;
  }
  return STORM_OK;
}

int locate_culverts(const Prep& S, int n, const double* pin, const double* pout,
                    std::vector<int>& cin, std::vector<int>& cout, std::vector<double>& rin,
                    std::string& err) {
  const int N = S.N;
  cin.assign(n, -1); cout.assign(n, -1); rin.assign(2 * (size_t)n, ZERO);
  // t_areas.f90:40-41 and pto_in_triangle.f90:20-43, cell by cell in input order
  auto inside = [&](int ii, double x, double y) {
    double vx[3], vy[3];
    for (int k = 0; k < 3; ++k) {
      const int vp = S.cvtx[(size_t)k * N + ii];
      vx[k] = S.vxyz[4 * (size_t)vp]; vy[k] = S.vxyz[4 * (size_t)vp + 1];
    }
    const double tarea = HALF * std::fabs((vx[0] - vx[1]) * (vy[0] + vy[1]) +
                                          (vx[1] - vx[2]) * (vy[1] + vy[2]) +
                                          (vx[2] - vx[0]) * (vy[2] + vy[0]));
    const double x1 = vx[0], y1 = vy[0], x2 = vx[1], y2 = vy[1], x3 = vx[2], y3 = vy[2];
    const double s = HALF / tarea * (y1 * x3 - x1 * y3 + (y3 - y1) * x + (x1 - x3) * y);
    const double t = HALF / tarea * (x1 * y2 - y1 * x2 + (y1 - y2) * x + (x2 - x1) * y);
    if (s < ZERO) return false;
    if (t < ZERO) return false;
    if (ONE - s - t < ZERO) return false;
    return true;
  };
  auto find = [&](double x, double y) {
    for (int j = 0; j < N; ++j)
      if (inside(S.iperm[j], x, y)) return j;
    return -1;
  };
  for (int m = 0; m < n; ++m) {
    cin[m] = find(pin[3 * m], pin[3 * m + 1]);
    if (cin[m] < 0) {
      err = "inlet of culvert " + std::to_string(m + 1) + " not in computational domain";
      return STORM_E_ARG;
    }
    const int ii = S.iperm[cin[m]];
    rin[2 * m] = pin[3 * m] - S.xc[ii];
    rin[2 * m + 1] = pin[3 * m + 1] - S.yc[ii];
    cout[m] = find(pout[2 * m], pout[2 * m + 1]);
    if (cout[m] < 0) {
      err = "outlet of culvert " + std::to_string(m + 1) + " not in computational domain";
      return STORM_E_ARG;
    }
  }
  return STORM_OK;
}

int build_prep(const storm_mesh_t* m, const storm_fvt_opts_t* o, Prep& S, const ReachSplit* split) {
  if (!m || !o) { S.err = "null mesh or options"; return STORM_E_ARG; }
  if (o->rkorder < 1 || o->rkorder > 3) { S.err = "Incorrect Runge-Kutta order"; return STORM_E_ARG; }
  if (o->flux < 0 || o->flux > 3 || o->gradient < 0 || o->gradient > 2 || o->limiter < 0 ||
      o->limiter > 2 || o->btype < 0 || o->btype > 1 || o->friction < 0 || o->friction > 1 ||
      o->dt_op < 0 || o->dt_op > 1 || o->stop_cr < 0 || o->stop_cr > 2 || o->fluxes < 0 ||
      o->fluxes > 1 || o->reorder < 0 || o->reorder > 2 || o->rsavg < 0 || o->rsavg > 1 || (o->rsavg && o->rsavg_n < 0) ||
      (o->rsavg && !(std::fabs(o->rsavg_eps) <= 1e300))) {
    S.err = "invalid option value"; return STORM_E_ARG;
  }
  // invisc_fluxes_xsf.f90 calls flux_ac only
  if (o->fluxes == 1 && o->flux != STORM_FLUX_AC) {
    S.err = "the xsf edge loop uses the Roe flux (flux_ac) only"; return STORM_E_ARG;
  }
  if (!(o->h_dry > 0) || !(o->h_wet >= 0) || !(o->cfl > 0)) { S.err = "invalid threshold or cfl"; return STORM_E_ARG; }
  if (m->n_pts < 3 || m->n_elems < 1 || !m->x || !m->y || !m->tri || !m->z || !m->h || !m->u ||
      !m->v || !m->cd) {
    S.err = "incomplete mesh description"; return STORM_E_ARG;
  }
  if (m->n_in < 0 || m->n_out < 0) { S.err = "negative string count"; return STORM_E_ARG; }
  const int N = m->n_elems, P = m->n_pts;
  // reaches: one unless a batch split is given
  ReachSplit one;
  if (!split) {
    one.cell_ptr = {0, N};
    one.in_ptr = {0, m->n_in};
    one.out_ptr = {0, m->n_out};
    split = &one;
  }
  const int NR = (int)split->cell_ptr.size() - 1;
  if (NR < 1 || split->cell_ptr[0] != 0 || split->cell_ptr[NR] != N ||
      (int)split->in_ptr.size() != NR + 1 || (int)split->out_ptr.size() != NR + 1 ||
      split->in_ptr[NR] != m->n_in || split->out_ptr[NR] != m->n_out) {
    S.err = "inconsistent reach split"; return STORM_E_ARG;
  }
  for (int r = 0; r < NR; ++r)
    if (split->cell_ptr[r + 1] <= split->cell_ptr[r] ||
        split->in_ptr[r + 1] - split->in_ptr[r] > MAX_STRINGS ||
        split->out_ptr[r + 1] - split->out_ptr[r] > MAX_STRINGS ||
        split->in_ptr[r + 1] < split->in_ptr[r] || split->out_ptr[r + 1] < split->out_ptr[r]) {
      S.err = "empty reach or too many boundary strings (max 64 of each kind per reach)";
      return STORM_E_ARG;
    }
  S.nr = NR;
  S.r_cell = split->cell_ptr;
  S.r_in = split->in_ptr;
  S.r_out = split->out_ptr;
  std::vector<int> cell_reach(N);
  for (int r = 0; r < NR; ++r)
    for (int i = split->cell_ptr[r]; i < split->cell_ptr[r + 1]; ++i) cell_reach[i] = r;
  for (int t = 0; t < 3 * N; ++t)
    if (m->tri[t] < 0 || m->tri[t] >= P) { S.err = "triangle vertex id out of range"; return STORM_E_MESH; }
  S.o = *o;
  S.N = N; S.P = P; S.R = o->rkorder;
  std::vector<int> ep, ee;
  int rcode = edge_table(m, ep, ee, S.err);
  if (rcode) return rcode;
  const int E = (int)(ep.size() / 2);
  S.E = E;
  const double* x = m->x;
  const double* y = m->y;
  const int32_t* tri = m->tri;

  // find_edges.f90:74-114: signed local edges
  std::vector<int> gedge(3 * (size_t)N, -1), gsign(3 * (size_t)N, 0);
  for (int j = 0; j < E; ++j)
    for (int kk = 0; kk < 2; ++kk) {
      int i = ee[2 * j + kk];
      if (i < 0) continue;
      for (int l = 0; l < 3; ++l) {
        int a = tri[3 * i + l], b = tri[3 * i + (l + 1) % 3];
        if (a == ep[2 * j] && b == ep[2 * j + 1]) { gedge[3 * i + l] = j; gsign[3 * i + l] = 1; break; }
        if (a == ep[2 * j + 1] && b == ep[2 * j]) { gedge[3 * i + l] = j; gsign[3 * i + l] = -1; break; }
      }
    }
  for (size_t t = 0; t < gedge.size(); ++t)
    if (gedge[t] < 0) { S.err = "edge data structure incomplete"; return STORM_E_MESH; }
  // edge_normals.f90:30-39 and find_edges.f90:159-163
  std::vector<double> nrm(2 * (size_t)E), tng(2 * (size_t)E), len(E);
  for (int k = 0; k < E; ++k) {
    double x1 = x[ep[2 * k]], y1 = y[ep[2 * k]], x2 = x[ep[2 * k + 1]], y2 = y[ep[2 * k + 1]];
    nrm[2 * k] = y2 - y1; nrm[2 * k + 1] = x1 - x2;
    tng[2 * k] = x2 - x1; tng[2 * k + 1] = y2 - y1;
    double dx = x1 - x2, dy = y1 - y2;
    len[k] = std::sqrt(dx * dx + dy * dy);
  }
  // t_areas.f90:33-52
  std::vector<double> area(N), perim(N), xc(N), yc(N);
  for (int i = 0; i < N; ++i) {
    int a = tri[3 * i], b = tri[3 * i + 1], d = tri[3 * i + 2];
    double x1 = x[a], y1 = y[a], x2 = x[b], y2 = y[b], x3 = x[d], y3 = y[d];
    area[i] = HALF * std::fabs((x1 - x2) * (y1 + y2) + (x2 - x3) * (y2 + y3) + (x3 - x1) * (y3 + y1));
    xc[i] = ONE_THIRD * (x1 + x2 + x3);
    yc[i] = ONE_THIRD * (y1 + y2 + y3);
    perim[i] = len[gedge[3 * i]] + len[gedge[3 * i + 1]] + len[gedge[3 * i + 2]];
    if (!(area[i] > 0)) { S.err = "degenerate triangle (zero area)"; return STORM_E_MESH; }
  }
  // boundary strings (ioflow.f90) and walls.f90:40-148
  S.n_in = m->n_in; S.n_out = m->n_out;
  std::vector<char> isbc(P, 0);
  const int nin_nodes = m->n_in ? m->in_ptr[m->n_in] : 0;
  const int nout_nodes = m->n_out ? m->out_ptr[m->n_out] : 0;
  for (int t = 0; t < nin_nodes; ++t) {
    if (m->in_nodes[t] < 0 || m->in_nodes[t] >= P) { S.err = "inflow node out of range"; return STORM_E_MESH; }
    isbc[m->in_nodes[t]] = 1;
  }
  for (int t = 0; t < nout_nodes; ++t) {
    if (m->out_nodes[t] < 0 || m->out_nodes[t] >= P) { S.err = "outflow node out of range"; return STORM_E_MESH; }
    isbc[m->out_nodes[t]] = 1;
  }
  std::vector<int> cnt(E, 0);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) cnt[gedge[3 * i + j]]++;
  std::vector<char> wall(E, 0);
  int nbpoly = 0, nwall1 = 0;
  for (int i = 0; i < E; ++i) {
    if (cnt[i] == 1) {
      nbpoly++;
      if (!(isbc[ep[2 * i]] && isbc[ep[2 * i + 1]])) { wall[i] = 1; nwall1++; }
    } else if (cnt[i] != 2) {
      S.err = "computational grid construction error, in edge assignment"; return STORM_E_MESH;
    }
  }
  std::unordered_map<uint64_t, int> emap;
  emap.reserve((size_t)E * 2);
  for (int k = E - 1; k >= 0; --k) emap[pair_key(ep[2 * k], ep[2 * k + 1])] = k;  // first match wins
  // ibc_arrays.f90:59-181,377-466: string edges, bcedges, qbar, hbar
  // edge -> (routine kind, string, position)
  std::vector<int> e_kind(E, -1), e_str(E, -1), e_pos(E, -1);
  std::vector<char> hbar_e(E, 0), qbar_e(E, 0);
  int nchain_edges = 0;
  S.in_type.assign(m->in_type, m->in_type + m->n_in);
  S.out_type.assign(m->out_type, m->out_type + m->n_out);
  auto chain_edges = [&](const int32_t* nodes, int a, int b, std::vector<int>& out) -> int {
    out.clear();
    for (int t = a; t + 1 < b; ++t) {
      auto it = emap.find(pair_key(nodes[t], nodes[t + 1]));
      if (it == emap.end()) {
        S.err = "boundary string nodes " + std::to_string(nodes[t]) + " " + std::to_string(nodes[t + 1]) +
                " are not connected by an edge";
        return STORM_E_MESH;
      }
      out.push_back(it->second);
    }
    return STORM_OK;
  };
  std::vector<int> ce;
  std::vector<std::vector<int>> in_edges(m->n_in), out_edges(m->n_out);
  for (int ch = 0; ch < m->n_in; ++ch) {
    int vt = m->in_type[ch];
    if (m->in_ptr[ch + 1] - m->in_ptr[ch] < 2) { S.err = "ERROR: n_qin <= 1."; return STORM_E_MESH; }
    if (chain_edges(m->in_nodes, m->in_ptr[ch], m->in_ptr[ch + 1], ce)) return STORM_E_MESH;
    in_edges[ch] = ce;
    nchain_edges += (int)ce.size();
    if ((vt == 1 || vt == 5) && ce.size() > 512) {
      S.err = "discharge (vtype 1/5) strings are limited to 512 edges"; return STORM_E_ARG;
    }
    int kind = (vt == 1) ? SK_IN_Q1 : (vt == 5) ? SK_IN_Q5 : (vt == 4) ? SK_IN_H : SK_NONE;
    if (vt < 0 || vt > 5) { S.err = "invalid value in vtype"; return STORM_E_ARG; }
    if (vt == 2) { S.err = "VTYPE = 2: option not implemented (enforce_bcs.f90:151-155)"; return STORM_E_ARG; }
    // enforce_bcs.f90:157-163 sets both velocity components of a vtype-3
    // string from two tables (timeserq, timeserq2); the mesh carries one
    if (vt == 3) { S.err = "VTYPE = 3 (two velocity tables) is not provided"; return STORM_E_ARG; }
    for (size_t p = 0; p < ce.size(); ++p) {
      if (cnt[ce[p]] != 1) { S.err = "inflow string runs through the interior of the mesh"; return STORM_E_MESH; }
      qbar_e[ce[p]] = 1;
      if (e_kind[ce[p]] < 0) { e_kind[ce[p]] = kind; e_str[ce[p]] = ch; e_pos[ce[p]] = (int)p; }
    }
  }
  for (int ch = 0; ch < m->n_out; ++ch) {
    int ht = m->out_type[ch];
    if (m->out_ptr[ch + 1] - m->out_ptr[ch] < 2) { S.err = "ERROR: n_hbc <= 1."; return STORM_E_MESH; }
    if (ht < 0 || ht > 2) { S.err = "invalid value in htype"; return STORM_E_ARG; }
    if (chain_edges(m->out_nodes, m->out_ptr[ch], m->out_ptr[ch + 1], ce)) return STORM_E_MESH;
    out_edges[ch] = ce;
    nchain_edges += (int)ce.size();
    int kind = (ht == 1) ? SK_OUT_H : (ht == 2) ? SK_OUT_CRIT : SK_OUT_FREE;
    for (size_t p = 0; p < ce.size(); ++p) {
      if (cnt[ce[p]] != 1) { S.err = "outflow string runs through the interior of the mesh"; return STORM_E_MESH; }
      hbar_e[ce[p]] = 1;
      if (e_kind[ce[p]] < 0) { e_kind[ce[p]] = kind; e_str[ce[p]] = ch; e_pos[ce[p]] = (int)p; }
    }
  }
  if (nbpoly - nwall1 != nchain_edges) {
    S.err = "number of boundary condition edges does not match number of boundary condition chains";
    return STORM_E_MESH;
  }
  for (int k = 0; k < E; ++k)
    if (cnt[k] == 1 && !wall[k] && e_kind[k] < 0) {
      S.err = "wall edges and inflow/outflow edges do not add up"; return STORM_E_MESH;
    }
  // enforce_bcs' tables (input ids; mapped to the internal order below).
  // ibc_arrays runs before dataprepFVT.f90:176-181 makes the edge normals
  // unit, so nrm still carries each edge's length here.
  auto edge_sign = [&](int k) {  // grid(t)%edge sign in the cell holding boundary edge k
    const int t = ee[2 * k];
    for (int l = 0; l < 3; ++l) if (gedge[3 * t + l] == k) return gsign[3 * t + l];
    return 0;
  };
  // ibc_arrays.f90:93-135: inflow string node normals, l_inflow
  std::vector<double> in_pn(2 * (size_t)nin_nodes, 0.0), in_len((size_t)nin_nodes, 0.0);
  for (int ch = 0; ch < m->n_in; ++ch) {
    const int a = m->in_ptr[ch], n = m->in_ptr[ch + 1] - a;
    double* pn = &in_pn[2 * (size_t)a];
    for (int i = 0; i + 1 < n; ++i) {
      const int k = in_edges[ch][i];
      const double sg = edge_sign(k) > 0 ? -1.0 : 1.0;
      pn[2 * i] = sg < 0 ? -nrm[2 * k] : nrm[2 * k];
      pn[2 * i + 1] = sg < 0 ? -nrm[2 * k + 1] : nrm[2 * k + 1];
      in_len[a + i] = len[k];
    }
    pn[2 * (n - 1)] = pn[2 * (n - 2)];
    pn[2 * (n - 1) + 1] = pn[2 * (n - 2) + 1];
    for (int i = n - 2; i >= 1; --i) {
      pn[2 * i] = pn[2 * i] + pn[2 * (i - 1)];
      pn[2 * i + 1] = pn[2 * i + 1] + pn[2 * (i - 1) + 1];
    }
    for (int i = 0; i < n; ++i) {
      const double l = std::sqrt(pn[2 * i] * pn[2 * i] + pn[2 * i + 1] * pn[2 * i + 1]);
      pn[2 * i] = pn[2 * i] / l;
      pn[2 * i + 1] = pn[2 * i + 1] / l;
    }
  }
  // walls.f90:150-176: wall points (walledg1 endpoints on no string)
  std::vector<int> wall_pts;
  {
    std::vector<char> seen(P, 0);
    for (int k = 0; k < E; ++k) {
      if (cnt[k] != 1 || !wall[k]) continue;
      for (int j = 0; j < 2; ++j) {
        const int p = ep[2 * k + j];
        if (!seen[p] && !isbc[p]) { seen[p] = 1; wall_pts.push_back(p); }
      }
    }
  }
  // ibc_arrays.f90:183-253,266-330: outflow string node normals and wall
  // tangents from the first two boundary-polygon edges through the node
  std::vector<std::vector<int>> node_bedges(P);
  for (int k = 0; k < E; ++k)
    if (cnt[k] == 1)
      for (int j = 0; j < 2; ++j) {
        std::vector<int>& be = node_bedges[ep[2 * k + j]];
        if (be.size() < 2) be.push_back(k);
      }
  auto node_normal = [&](int p, double& nx, double& ny) -> bool {
    const std::vector<int>& be = node_bedges[p];
    if (be.empty()) return false;
    const int e1 = be[0], e2 = be.size() > 1 ? be[1] : be[0];
    const double s1 = edge_sign(e1) > 0 ? 1.0 : -1.0, s2 = edge_sign(e2) > 0 ? 1.0 : -1.0;
    const double n1x = s1 > 0 ? nrm[2 * e1] : -nrm[2 * e1], n1y = s1 > 0 ? nrm[2 * e1 + 1] : -nrm[2 * e1 + 1];
    const double n2x = s2 > 0 ? nrm[2 * e2] : -nrm[2 * e2], n2y = s2 > 0 ? nrm[2 * e2 + 1] : -nrm[2 * e2 + 1];
    nx = n1x + n2x;
    ny = n1y + n2y;
    const double l = std::sqrt(nx * nx + ny * ny);
    nx = nx / l;
    ny = ny / l;
    return true;
  };
  std::vector<int> on_vtx;          // (vertex, normal) entries of the velocity check
  std::vector<double> on_nrm;
  for (int ch = 0; ch < m->n_out; ++ch) {
    const int ht = m->out_type[ch];
    const bool check = ht == 0 || (ht == 1 && o->hvalve);
    for (int t = m->out_ptr[ch]; t < m->out_ptr[ch + 1]; ++t) {
      double nx, ny;
      if (!node_normal(m->out_nodes[t], nx, ny)) {
        S.err = "outflow string node on no boundary edge"; return STORM_E_MESH;
      }
      if (!check) continue;
      on_vtx.push_back(m->out_nodes[t]);
      on_nrm.push_back(nx);
      on_nrm.push_back(ny);
    }
  }
  std::vector<double> wall_tg(2 * wall_pts.size());
  for (size_t w = 0; w < wall_pts.size(); ++w) {
    double nx, ny;
    if (!node_normal(wall_pts[w], nx, ny)) { S.err = "wall point on no boundary edge"; return STORM_E_MESH; }
    wall_tg[2 * w] = -ny;
    wall_tg[2 * w + 1] = nx;
  }
  node_bedges.clear();
  node_bedges.shrink_to_fit();
  // time tables
  S.qts_ptr.assign(m->n_in + 1, 0);
  S.hts_ptr.assign(m->n_out + 1, 0);
  if (m->n_in) {
    S.qts_ptr.assign(m->qts_ptr, m->qts_ptr + m->n_in + 1);
    S.qts_t.assign(m->qts_t, m->qts_t + S.qts_ptr[m->n_in]);
    S.qts_v.assign(m->qts_v, m->qts_v + S.qts_ptr[m->n_in]);
  }
  if (m->n_out) {
    S.hts_ptr.assign(m->hts_ptr, m->hts_ptr + m->n_out + 1);
    S.hts_t.assign(m->hts_t, m->hts_t + S.hts_ptr[m->n_out]);
    S.hts_v.assign(m->hts_v, m->hts_v + S.hts_ptr[m->n_out]);
  }
  // node_db.f90:60-68: vertex -> cells, ascending cell ids
  std::vector<int> n2t_ptr(P + 1, 0), n2t;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) n2t_ptr[tri[3 * i + j] + 1]++;
  for (int p = 0; p < P; ++p) n2t_ptr[p + 1] += n2t_ptr[p];
  n2t.resize(n2t_ptr[P]);
  {
    std::vector<int> fill(P, 0);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < 3; ++j) {
        int p = tri[3 * i + j];
        n2t[n2t_ptr[p] + fill[p]++] = i;
      }
  }
  // elemnt_db.f90:42-128 (molecule) and :226-233 (edge neighbours)
  // with ACTIVATE HDLS the gradients use t2tHD instead (cell_grads.f90:37-42)
  std::vector<int> t2t_ptr, t2t;
  if (o->gradient == STORM_GRAD_HDLS) {
    std::vector<int> base(N);
    for (int i = 0; i < N; ++i) base[i] = split->cell_ptr[cell_reach[i]];
    rcode = hd_molecule(N, tri, n2t_ptr, n2t, base, t2t_ptr, t2t, S.err);
  } else {
    rcode = molecule(N, E, ee, t2t_ptr, t2t, S.err);
  }
  if (rcode) return rcode;
  std::vector<int> t2t3(3 * (size_t)N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) {
      int k = gedge[3 * i + j];
      int e1 = ee[2 * k];
      if (e1 == i) e1 = ee[2 * k + 1];
      t2t3[3 * i + j] = e1;
    }
  // c2v_setup.f90:106-119 (type 3)
  std::vector<double> c2vw(n2t.size()), c2vwid(n2t.size());
  for (int i = 0; i < P; ++i) {
    double tdist = ZERO;
    for (int j = n2t_ptr[i]; j < n2t_ptr[i + 1]; ++j) {
      int k = n2t[j];
      double xk = xc[k] - x[i], yk = yc[k] - y[i];
      c2vw[j] = ONE / std::sqrt(xk * xk + yk * yk);
      c2vwid[j] = c2vw[j];
      tdist = tdist + c2vw[j];
    }
    for (int j = n2t_ptr[i]; j < n2t_ptr[i + 1]; ++j) c2vw[j] = c2vw[j] / tdist;
  }
  auto vtx2ctr = [&](const double* va, std::vector<double>& ca) {  // vtx2ctr.f90:32-35
    ca.resize(N);
    for (int i = 0; i < N; ++i)
      ca[i] = (va[tri[3 * i]] + va[tri[3 * i + 1]] + va[tri[3 * i + 2]]) * ONE_THIRD;
  };
  auto ctr2vtx = [&](const double* ca, std::vector<double>& va) {  // ctr2vtx.f90:73-80
    va.resize(P);
    for (int i = 0; i < P; ++i) {
      double a = ZERO;
      for (int j = n2t_ptr[i]; j < n2t_ptr[i + 1]; ++j) a = a + c2vw[j] * ca[n2t[j]];
      va[i] = a;
    }
  };
  // dataprepFVT.f90:73-140
  S.vin_h.clear(); S.vin_u.clear(); S.vin_v.clear(); S.vin_cd.clear();
  if (m->h_at_vtx) S.vin_h.assign(m->h, m->h + P);
  if (m->u_at_vtx) S.vin_u.assign(m->u, m->u + P);
  if (m->v_at_vtx) S.vin_v.assign(m->v, m->v + P);
  if (m->cd_at_vtx) S.vin_cd.assign(m->cd, m->cd + P);
  std::vector<double> z, zvtx, h, u, v, cd, tmp;
  if (m->z_at_vtx) { vtx2ctr(m->z, z); zvtx.assign(m->z, m->z + P); }
  else { ctr2vtx(m->z, zvtx); vtx2ctr(zvtx.data(), z); }
  if (m->h_at_vtx) vtx2ctr(m->h, h); else h.assign(m->h, m->h + N);
  if (m->u_at_vtx) vtx2ctr(m->u, u); else u.assign(m->u, m->u + N);
  if (m->v_at_vtx) vtx2ctr(m->v, v); else v.assign(m->v, m->v + N);
  if (m->cd_at_vtx) vtx2ctr(m->cd, cd); else cd.assign(m->cd, m->cd + N);
  // dataprepFVT.f90:164-181: rc vectors; unit normals and tangents
  std::vector<double> rcx(3 * (size_t)N), rcy(3 * (size_t)N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < 3; ++j) {
      int k = gedge[3 * i + j];
      double xm = HALF * (x[ep[2 * k]] + x[ep[2 * k + 1]]);
      double ym = HALF * (y[ep[2 * k]] + y[ep[2 * k + 1]]);
      rcx[3 * i + j] = xm - xc[i];
      rcy[3 * i + j] = ym - yc[i];
    }
  for (int k = 0; k < E; ++k) {
    nrm[2 * k] = nrm[2 * k] / len[k]; nrm[2 * k + 1] = nrm[2 * k + 1] / len[k];
    tng[2 * k] = tng[2 * k] / len[k]; tng[2 * k + 1] = tng[2 * k + 1] / len[k];
  }
  // lsq_setup.f90:105-145 (weighted, type 2), or hdlsq_setup.f90:39-74
  // (unweighted, type 1) on t2tHD
  const bool hd = o->gradient == STORM_GRAD_HDLS;
  std::vector<double> mdx(t2t.size()), mdy(t2t.size()), mw(t2t.size()), ai(N), bi(N), ci(N), det(N);
  for (int i = 0; i < N; ++i) {
    double a = ZERO, b = ZERO, c = ZERO;
    for (int t = t2t_ptr[i]; t < t2t_ptr[i + 1]; ++t) {
      int l = t2t[t];
      mdx[t] = xc[l] - xc[i];
      mdy[t] = yc[l] - yc[i];
      mw[t] = hd ? ONE : ONE / (mdx[t] * mdx[t] + mdy[t] * mdy[t]);
    }
    for (int t = t2t_ptr[i]; t < t2t_ptr[i + 1]; ++t) {
      a = a + mw[t] * mdx[t] * mdx[t];
      b = b + mw[t] * mdx[t] * mdy[t];
      c = c + mw[t] * mdy[t] * mdy[t];
    }
    ai[i] = a; bi[i] = b; ci[i] = c;
    det[i] = a * c - b * b;
  }
  // Green-Gauss gradient coefficients on the triangle of neighbour centroids
  // (gaussgrad_setup.f90:33-81 and ghost_cell.f90:34-41, corrected: the
  // neighbour's centroid, t2t3(j,i), outward normals of the auxiliary triangle)
  std::vector<int> gconn(3 * (size_t)N);
  std::vector<double> gcx(3 * (size_t)N), gcy(3 * (size_t)N), gwt(3 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    double xp[3], yp[3];
    for (int j = 0; j < 3; ++j) {
      int k = t2t3[3 * i + j];
      if (k >= 0) { xp[j] = xc[k]; yp[j] = yc[k]; gconn[3 * i + j] = k; }
      else {
        int l = gedge[3 * i + j];
        double x1 = x[ep[2 * l]], y1 = y[ep[2 * l]], x2 = x[ep[2 * l + 1]], y2 = y[ep[2 * l + 1]];
        double ax = xc[i] - x1, ay = yc[i] - y1, bx = x2 - x1, by = y2 - y1;
        double den = bx * bx + by * by;
        double cx = (bx * (ax * bx + ay * by) + by * (ay * bx - ax * by)) / den;
        double cy = (bx * (ax * by - ay * bx) + by * (ax * bx + ay * by)) / den;
        xp[j] = cx + x1; yp[j] = cy + y1; gconn[3 * i + j] = i;
      }
    }
    double ar = HALF * ((xp[1] - xp[0]) * (yp[2] - yp[0]) - (xp[2] - xp[0]) * (yp[1] - yp[0]));
    double n1x = yp[1] - yp[0], n1y = xp[0] - xp[1];
    double n2x = yp[2] - yp[1], n2y = xp[1] - xp[2];
    double n3x = yp[0] - yp[2], n3y = xp[2] - xp[0];
    gcx[3 * i] = HALF * (n1x + n3x) / ar; gcx[3 * i + 1] = HALF * (n1x + n2x) / ar; gcx[3 * i + 2] = HALF * (n2x + n3x) / ar;
    gcy[3 * i] = HALF * (n1y + n3y) / ar; gcy[3 * i + 1] = HALF * (n1y + n2y) / ar; gcy[3 * i + 2] = HALF * (n2y + n3y) / ar;
    double den = (xp[1] - xp[0]) * (yp[2] - yp[1]) - (xp[2] - xp[1]) * (yp[1] - yp[0]);
    gwt[3 * i] = ((xp[1] - xc[i]) * (yp[2] - yp[1]) - (xp[2] - xp[1]) * (yp[1] - yc[i])) / den;
    gwt[3 * i + 1] = ((xp[2] - xc[i]) * (yp[0] - yp[2]) - (xp[0] - xp[2]) * (yp[2] - yc[i])) / den;
    gwt[3 * i + 2] = ((xp[0] - xc[i]) * (yp[1] - yp[0]) - (xp[1] - xp[0]) * (yp[0] - yc[i])) / den;
  }
  // dataprepFVT.f90:278-286, per reach
  S.r_bch.assign(NR, ZERO);
  for (int k = 0; k < E; ++k)
    if (qbar_e[k] && cnt[k] == 1) {
      double& bm = S.r_bch[cell_reach[ee[2 * k]]];
      bm = std::max(bm, zvtx[ep[2 * k]]);
      bm = std::max(bm, zvtx[ep[2 * k + 1]]);
    }
  // dataprepFVT.f90:293-306: highest vertex; bed_slope_VTX.f90:22-38
  std::vector<double> zhigh(N), dzx(N), dzy(N);
  for (int i = 0; i < N; ++i) {
    int p[3] = {tri[3 * i], tri[3 * i + 1], tri[3 * i + 2]};
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k)
        if (zvtx[p[k]] > zvtx[p[k + 1]]) std::swap(p[k], p[k + 1]);
    zhigh[i] = zvtx[p[2]];
    int a = tri[3 * i], b = tri[3 * i + 1], c = tri[3 * i + 2];
    double x1 = x[a], y1 = y[a], x2 = x[b], y2 = y[b], x3 = x[c], y3 = y[c];
    double z1 = zvtx[a], z2 = zvtx[b], z3 = zvtx[c];
    dzx[i] = HALF * (z1 * (y2 - y3) + z2 * (y3 - y1) + z3 * (y1 - y2)) / area[i];
    dzy[i] = HALF * (z1 * (x3 - x2) + z2 * (x1 - x3) + z3 * (x2 - x1)) / area[i];
  }

  // consistent orientation: every interior edge runs opposite ways in its two cells
  for (int k = 0; k < E; ++k) {
    if (cnt[k] != 2) continue;
    int e1 = ee[2 * k], e2 = ee[2 * k + 1], s1 = 0, s2 = 0;
    for (int l = 0; l < 3; ++l) {
      if (gedge[3 * e1 + l] == k) s1 = gsign[3 * e1 + l];
      if (gedge[3 * e2 + l] == k) s2 = gsign[3 * e2 + l];
    }
    if (s1 != -s2) { S.err = "inconsistently oriented triangles"; return STORM_E_MESH; }
  }

  // ---- cell ordering --------------------------------------------------
  std::vector<int> order;
  order.resize(N);
  std::iota(order.begin(), order.end(), 0);
  if (o->reorder == 1)
    for (int r = 0; r < NR; ++r) {
      const int a = split->cell_ptr[r], n = split->cell_ptr[r + 1] - a;
      std::vector<double> rx(xc.begin() + a, xc.begin() + a + n), ry(yc.begin() + a, yc.begin() + a + n);
      const std::vector<int> ro = hilbert_order(n, rx, ry);
      for (int k = 0; k < n; ++k) order[a + k] = a + ro[k];
    }
  else if (o->reorder == 2)
    for (int r = 0; r < NR; ++r) {
      const int a = split->cell_ptr[r], n = split->cell_ptr[r + 1] - a;
      const std::vector<int> ro = rcm_order(a, n, t2t3);
      for (int k = 0; k < n; ++k) order[a + k] = a + ro[k];
    }
  S.perm = order;
  S.iperm.assign(N, 0);
  for (int i = 0; i < N; ++i) S.iperm[order[i]] = i;
  // vertices in order of first touch by the ordered cells
  S.vperm.clear(); S.viperm.assign(P, -1);
  for (int ii = 0; ii < N; ++ii) {
    int i = order[ii];
    for (int j = 0; j < 3; ++j) {
      int p = tri[3 * i + j];
      if (S.viperm[p] < 0) { S.viperm[p] = (int)S.vperm.size(); S.vperm.push_back(p); }
    }
  }
  for (int p = 0; p < P; ++p)
    if (S.viperm[p] < 0) { S.viperm[p] = (int)S.vperm.size(); S.vperm.push_back(p); }
  // Within each run of 64 vertices (one wave of the vertex pass), vertices
  // by degree.  The vertex pass reads row r of its incident-cell and weight
  // tables only for vertices of degree > r, but a cache line of a row is
  // fetched when any of its vertices needs it; neighbours of like degree
  // share lines, and the rows past a line's largest degree are not read.
  for (int b = 0; b < P; b += 64) {
    const int e = std::min(P, b + 64);
    std::stable_sort(S.vperm.begin() + b, S.vperm.begin() + e, [&](int a, int c) {
      return n2t_ptr[a + 1] - n2t_ptr[a] < n2t_ptr[c + 1] - n2t_ptr[c];
    });
  }
  for (int pp = 0; pp < P; ++pp) S.viperm[S.vperm[pp]] = pp;
  const std::vector<int>& ip = S.iperm;
  const std::vector<int>& vip = S.viperm;
  const size_t N3 = 3 * (size_t)N;
  S.cvtx.assign(N3, 0); S.cnbr.assign(N3, -1); S.cmeta.assign(N, 0);
  S.xc.resize(N); S.yc.resize(N); S.z.resize(N); S.cd.resize(N);
  S.gconn.assign(N3, 0); S.gcx.assign(N3, 0); S.gcy.assign(N3, 0); S.gwt.assign(N3, 0);
  S.mptr.assign(N + 1, 0);
  S.h0.resize(N); S.u0.resize(N); S.v0.resize(N);
  for (int ii = 0; ii < N; ++ii) S.mptr[ii + 1] = S.mptr[ii] + (t2t_ptr[order[ii] + 1] - t2t_ptr[order[ii]]);
  S.mcell.resize(S.mptr[N]);
  for (int ii = 0; ii < N; ++ii) {
    const int i = order[ii];
    S.xc[ii] = xc[i]; S.yc[ii] = yc[i]; S.z[ii] = z[i]; S.cd[ii] = cd[i];
    S.h0[ii] = h[i]; S.u0[ii] = u[i]; S.v0[ii] = v[i];
    for (int t = 0; t < t2t_ptr[i + 1] - t2t_ptr[i]; ++t)
      S.mcell[S.mptr[ii] + t] = ip[t2t[t2t_ptr[i] + t]];
    uint64_t keys[3];
    uint64_t meta = 0;
    for (int j = 0; j < 3; ++j) {
      const size_t s = (size_t)j * N + ii;
      const int k = gedge[3 * i + j];
      S.cvtx[s] = vip[tri[3 * i + j]];
      S.gconn[s] = ip[gconn[3 * i + j]];
      S.gcx[s] = gcx[3 * i + j]; S.gcy[s] = gcy[3 * i + j]; S.gwt[s] = gwt[3 * i + j];
      const int e1 = ee[2 * k];
      uint64_t slot = 0;
      if (gsign[3 * i + j] < 0) slot |= 1ull << CMS_REV;
      if (cnt[k] == 2) {
        int nb = t2t3[3 * i + j];
        S.cnbr[s] = ip[nb];
        int nl = -1;
        for (int l = 0; l < 3; ++l) if (gedge[3 * nb + l] == k) { nl = l; break; }
        slot |= (uint64_t)SK_INTERIOR | ((uint64_t)nl << CMS_NL);
        if (e1 == i) slot |= 1ull << CMS_LEFT;
        keys[j] = ((uint64_t)0 << 60) | (uint64_t)k;
      } else if (wall[k]) {
        slot |= (uint64_t)SK_WALL | (1ull << CMS_LEFT);
        keys[j] = ((uint64_t)1 << 60) | (uint64_t)k;
      } else {
        uint64_t kind = (uint64_t)e_kind[k];
        // reference routine order: inflowQbyH (vtype 1 strings, then 5),
        // inflowH, outflowbyH, crit_outflow, free_flow
        uint64_t rank = (kind == SK_IN_Q1) ? 0 : (kind == SK_IN_Q5) ? 1 : (kind == SK_IN_H) ? 2
                      : (kind == SK_OUT_H) ? 3 : (kind == SK_OUT_CRIT) ? 4 : (kind == SK_OUT_FREE) ? 5 : 6;
        const int rr = cell_reach[i];
        const int base = (kind == SK_IN_Q1 || kind == SK_IN_Q5 || kind == SK_IN_H || kind == SK_NONE)
                             ? split->in_ptr[rr] : split->out_ptr[rr];
        slot |= kind | (1ull << CMS_LEFT) | ((uint64_t)(e_str[k] - base) << CMS_CHAIN);
        keys[j] = ((uint64_t)2 << 60) | (rank << 52) | ((uint64_t)e_str[k] << 32) | (uint64_t)e_pos[k];
      }
      if (hbar_e[k] && cnt[k] == 1) slot |= 1ull << CMS_HBAR;
      meta |= slot << (CM_SLOT0 + CM_SLOT_BITS * j);
    }
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int b) { return keys[a] < keys[b]; });
    meta |= (uint64_t)(ord[0] | (ord[1] << 2) | (ord[2] << 4));
    S.cmeta[ii] = meta;
  }
  // vertices: coordinates and bed, vertex -> cells in input cell order
  S.vptr.assign(P + 1, 0);
  for (int pp = 0; pp < P; ++pp) {
    int p = S.vperm[pp];
    S.vptr[pp + 1] = S.vptr[pp] + (n2t_ptr[p + 1] - n2t_ptr[p]);
  }
  S.vcell.resize(S.vptr[P]); S.vbc.assign(P, -1); S.vxyz.assign(4 * (size_t)P, 0.0);
  for (int pp = 0; pp < P; ++pp) {
    int p = S.vperm[pp];
    S.vxyz[4 * (size_t)pp] = x[p];
    S.vxyz[4 * (size_t)pp + 1] = y[p];
    S.vxyz[4 * (size_t)pp + 2] = zvtx[p];
    for (int t = 0; t < n2t_ptr[p + 1] - n2t_ptr[p]; ++t)
      S.vcell[S.vptr[pp] + t] = ip[n2t[n2t_ptr[p] + t]];
  }
  // enforce_bcs.f90:40-56: nodes of stage strings (last string wins, as the
  // reference loops over strings in order)
  for (int ch = 0; ch < m->n_out; ++ch) {
    if (m->out_type[ch] != 1) continue;
    for (int t = m->out_ptr[ch]; t < m->out_ptr[ch + 1]; ++t) S.vbc[vip[m->out_nodes[t]]] = ch;
  }
  // the vertex pass's tables: row r of vertex p at r*P + p (n2t order), the
  // inverse-distance weights of c2v_setup.f90:106-113, and per vertex
  // {zvtx, string index | degree << 32 | reach << 40}
  S.vW = 0;
  for (int p = 0; p < P; ++p) S.vW = std::max(S.vW, n2t_ptr[p + 1] - n2t_ptr[p]);
  S.vell.assign((size_t)S.vW * P, -1);
  S.vwt.assign((size_t)S.vW * P, 0.0);
  S.vzb.assign(2 * (size_t)P, 0.0);
  for (int pp = 0; pp < P; ++pp) {
    const int p = S.vperm[pp], deg = n2t_ptr[p + 1] - n2t_ptr[p];
    for (int t = 0; t < deg; ++t) {
      S.vell[(size_t)t * P + pp] = ip[n2t[n2t_ptr[p] + t]];
      S.vwt[(size_t)t * P + pp] = c2vwid[n2t_ptr[p] + t];
    }
    if (deg > 255) { S.err = "a vertex has more than 255 triangles"; return STORM_E_MESH; }
    const int vr = deg > 0 ? cell_reach[n2t[n2t_ptr[p]]] : 0;
    const uint64_t bits = (uint64_t)(uint32_t)S.vbc[pp] | ((uint64_t)deg << 32) | ((uint64_t)vr << 40);
    S.vzb[2 * (size_t)pp] = zvtx[p];
    std::memcpy(&S.vzb[2 * (size_t)pp + 1], &bits, sizeof(bits));
  }
  // node_db.f90:165-185: n2n in the internal vertex order, each row in the
  // order the edge loop fills it
  {
    std::vector<int> cnt_in(P + 1, 0), row_in(2 * (size_t)E), fill(P, 0);
    for (int k = 0; k < E; ++k) { cnt_in[ep[2 * k] + 1]++; cnt_in[ep[2 * k + 1] + 1]++; }
    for (int p = 0; p < P; ++p) cnt_in[p + 1] += cnt_in[p];
    for (int k = 0; k < E; ++k) {
      const int p1 = ep[2 * k], p2 = ep[2 * k + 1];
      row_in[cnt_in[p1] + fill[p1]++] = p2;
      row_in[cnt_in[p2] + fill[p2]++] = p1;
    }
    S.nn_ptr.assign(P + 1, 0);
    for (int pp = 0; pp < P; ++pp) {
      const int p = S.vperm[pp];
      S.nn_ptr[pp + 1] = S.nn_ptr[pp] + (cnt_in[p + 1] - cnt_in[p]);
    }
    S.nn_idx.resize(2 * (size_t)E);
    for (int pp = 0; pp < P; ++pp) {
      const int p = S.vperm[pp];
      for (int t = cnt_in[p]; t < cnt_in[p + 1]; ++t)
        S.nn_idx[S.nn_ptr[pp] + (t - cnt_in[p])] = vip[row_in[t]];
    }
  }
  // enforce_bcs' tables in the internal vertex order
  S.vo_ptr.assign(P + 1, 0);
  for (int v : on_vtx) S.vo_ptr[vip[v] + 1]++;
  for (int pp = 0; pp < P; ++pp) S.vo_ptr[pp + 1] += S.vo_ptr[pp];
  S.vo_nrm.assign(on_nrm.size(), 0.0);
  {
    std::vector<int> fill(S.vo_ptr.begin(), S.vo_ptr.end() - 1);
    for (size_t e = 0; e < on_vtx.size(); ++e) {
      const int at = fill[vip[on_vtx[e]]]++;
      S.vo_nrm[2 * (size_t)at] = on_nrm[2 * e];
      S.vo_nrm[2 * (size_t)at + 1] = on_nrm[2 * e + 1];
    }
  }
  S.vwall.assign(P, -1);
  S.wtang = wall_tg;
  for (size_t w = 0; w < wall_pts.size(); ++w) S.vwall[vip[wall_pts[w]]] = (int)w;
  S.is_ptr.assign(m->n_in + 1, 0);
  S.is_node.resize(nin_nodes);
  for (int ch = 0; ch < m->n_in; ++ch) S.is_ptr[ch + 1] = m->in_ptr[ch + 1];
  for (int t = 0; t < nin_nodes; ++t) S.is_node[t] = vip[m->in_nodes[t]];
  S.is_pn = in_pn;
  S.is_len = in_len;
  // inflowQbyH strings: (cell, local edge) in string order, with the edge's
  // rc vector, unit normal (as the reference orients it) and length
  S.qs_ptr.assign(m->n_in + 1, 0);
  S.qs_cell.clear(); S.qs_slot.clear();
  S.qs_rcx.clear(); S.qs_rcy.clear(); S.qs_nx.clear(); S.qs_ny.clear(); S.qs_len.clear();
  for (int ch = 0; ch < m->n_in; ++ch) {
    int vt = m->in_type[ch];
    if (vt == 1 || vt == 5) {
      for (int k : in_edges[ch]) {
        int i = ee[2 * k];
        int l = -1;
        for (int t = 0; t < 3; ++t) if (gedge[3 * i + t] == k) { l = t; break; }
        S.qs_cell.push_back(ip[i]);
        S.qs_slot.push_back(l);
        S.qs_rcx.push_back(rcx[3 * i + l]);
        S.qs_rcy.push_back(rcy[3 * i + l]);
        const double sg = (gsign[3 * i + l] < 0) ? -1.0 : 1.0;
        S.qs_nx.push_back(sg < 0 ? -nrm[2 * k] : nrm[2 * k]);
        S.qs_ny.push_back(sg < 0 ? -nrm[2 * k + 1] : nrm[2 * k + 1]);
        S.qs_len.push_back(len[k]);
      }
    }
    S.qs_ptr[ch + 1] = (int)S.qs_cell.size();
  }
  S.r_last.resize(NR);
  for (int r = 0; r < NR; ++r) S.r_last[r] = ip[split->cell_ptr[r + 1] - 1];
  {
    std::vector<std::vector<int>> sv(NR);
    for (int pp = 0; pp < P; ++pp) {
      if (S.vbc[pp] < 0) continue;
      const int p = S.vperm[pp];
      const int deg = n2t_ptr[p + 1] - n2t_ptr[p];
      const int vr = deg > 0 ? cell_reach[n2t[n2t_ptr[p]]] : 0;
      sv[vr].push_back(pp);
    }
    S.r_svptr.assign(1, 0); S.r_svtx.clear();
    for (int r = 0; r < NR; ++r) {
      S.r_svtx.insert(S.r_svtx.end(), sv[r].begin(), sv[r].end());
      S.r_svptr.push_back((int)S.r_svtx.size());
    }
  }
  S.rs_lptr.clear(); S.rs_cell.clear(); S.rs_nptr.clear(); S.rs_nbr.clear(); S.r_lev.clear();
  if (o->rsavg) {
    // the smoothing reads elemnt_db's t2t whatever the gradients use
    std::vector<int> ep2, et2;
    if (o->gradient == STORM_GRAD_HDLS) {
      rcode = molecule(N, E, ee, ep2, et2, S.err);
      if (rcode) return rcode;
    }
    const std::vector<int>& rp = (o->gradient == STORM_GRAD_HDLS) ? ep2 : t2t_ptr;
    const std::vector<int>& rt = (o->gradient == STORM_GRAD_HDLS) ? et2 : t2t;
    rsavg_schedule(S, rp, rt, split->cell_ptr);
    rsavg_fast_tables(S);
  }
  return STORM_OK;
}

void dataprep_fields(const Prep& S, double* z, double* zvtx, double* zeta, double* cd,
                     double* cdvtx, double* hvtx, double* uvtx, double* vvtx, double* zetavtx) {
  const int N = S.N, P = S.P;
  const size_t PP = (size_t)P;
  // ctr2vtx type 3 of a cell field (internal order) at every vertex, written
  // in input vertex order
  auto ctr2vtx = [&](const std::vector<double>& f, double* out) {
    for (int pp = 0; pp < P; ++pp) {
      const int deg = S.vptr[pp + 1] - S.vptr[pp];
      double tdist = ZERO;
      for (int t = 0; t < deg; ++t) tdist = tdist + S.vwt[(size_t)t * PP + pp];
      double a = ZERO;
      for (int t = 0; t < deg; ++t)
        a = a + (S.vwt[(size_t)t * PP + pp] / tdist) * f[S.vell[(size_t)t * PP + pp]];
      out[S.vperm[pp]] = a;
    }
  };
  auto cells_out = [&](const std::vector<double>& f, double* out) {
    for (int ii = 0; ii < N; ++ii) out[S.perm[ii]] = f[ii];
  };
  std::vector<double> zv(P);  // zvtx, input order
  for (int pp = 0; pp < P; ++pp) zv[S.vperm[pp]] = S.vxyz[4 * (size_t)pp + 2];
  if (z) cells_out(S.z, z);
  if (zvtx) std::copy(zv.begin(), zv.end(), zvtx);
  if (cd) cells_out(S.cd, cd);
  if (cdvtx) {
    if (!S.vin_cd.empty()) std::copy(S.vin_cd.begin(), S.vin_cd.end(), cdvtx);
    else ctr2vtx(S.cd, cdvtx);
  }
  if (uvtx) {
    if (!S.vin_u.empty()) std::copy(S.vin_u.begin(), S.vin_u.end(), uvtx);
    else ctr2vtx(S.u0, uvtx);
  }
  if (vvtx) {
    if (!S.vin_v.empty()) std::copy(S.vin_v.begin(), S.vin_v.end(), vvtx);
    else ctr2vtx(S.v0, vvtx);
  }
  if (!zeta && !hvtx && !zetavtx) return;
  std::vector<double> zc(N), zvt(P);  // zeta (internal order), zetavtx (input order)
  if (!S.vin_h.empty()) {
    // dataprepFVT.f90:90-96
    for (int p = 0; p < P; ++p) zvt[p] = S.vin_h[p] + zv[p];
    for (int ii = 0; ii < N; ++ii) {
      const int* cv = &S.cvtx[0];
      const double a = zvt[S.vperm[cv[ii]]], b = zvt[S.vperm[cv[(size_t)N + ii]]],
                   c = zvt[S.vperm[cv[2 * (size_t)N + ii]]];
      zc[ii] = (a + b + c) * ONE_THIRD;
    }
    if (hvtx) std::copy(S.vin_h.begin(), S.vin_h.end(), hvtx);
  } else {
    // dataprepFVT.f90:97-104
    for (int ii = 0; ii < N; ++ii) zc[ii] = S.z[ii] + S.h0[ii];
    ctr2vtx(zc, zvt.data());
    if (hvtx)
      for (int p = 0; p < P; ++p) hvtx[p] = zvt[p] - zv[p];
  }
  if (zeta) cells_out(zc, zeta);
  if (zetavtx) std::copy(zvt.begin(), zvt.end(), zetavtx);
}

int build_tiles(Prep& S, int cap) {
  const int N = S.N;
  if (cap < 16 || cap > 0xFFFF) { S.err = "bad tile capacity"; return STORM_E_ARG; }
  const int* cn = S.cnbr.data();
  // greedy runs of consecutive cells: grow while the run and its halo of
  // edge neighbours fit the capacity
  std::vector<int> hstamp(N, -1);
  S.t_ptr.assign(1, 0);
  S.r_tile.assign(1, 0);
  int start = 0, tile = 0, reach = 0;
  while (start < N) {
    while (S.r_cell[reach + 1] <= start) {
      S.r_tile.push_back(tile);
      ++reach;
    }
    const int end = S.r_cell[reach + 1];
    int T = 0, H = 0;
    while (start + T < end) {
      const int c = start + T;
      int newh[3], nnh = 0;
      const bool was_halo = hstamp[c] == tile;
      for (int j = 0; j < 3; ++j) {
        const int nb = cn[(size_t)j * N + c];
        if (nb < 0 || (nb >= start && nb <= c) || hstamp[nb] == tile) continue;
        bool dup = false;
        for (int k = 0; k < nnh; ++k) dup = dup || newh[k] == nb;
        if (!dup) newh[nnh++] = nb;
      }
      const int T2 = T + 1, H2 = H + nnh - (was_halo ? 1 : 0);
      if (T2 + H2 > cap || T2 > tile_tmax(cap)) break;
      for (int k = 0; k < nnh; ++k) hstamp[newh[k]] = tile;
      if (was_halo) hstamp[c] = -1;
      T = T2; H = H2;
    }
    if (T == 0) { S.err = "tile capacity too small for the mesh"; return STORM_E_ARG; }
    start += T;
    S.t_ptr.push_back(start);
    ++tile;
  }
  while ((int)S.r_tile.size() < S.nr + 1) S.r_tile.push_back(tile);
  const int nt = tile;
  S.nt = nt;
  S.cap = cap;
  std::vector<int> tile_reach(nt);
  for (int r = 0; r < S.nr; ++r)
    for (int b = S.r_tile[r]; b < S.r_tile[r + 1]; ++b) tile_reach[b] = r;
  // per tile: halo list (ascending ids) and 16-bit local neighbour ids
  S.t_hptr.assign(1, 0); S.t_hcell.clear();
  S.t_cln.assign(N, 0);
  S.t_qmask.assign(nt, 0);
  std::vector<int> hloc(N, -1);
  std::vector<int> hl;
  for (int b = 0; b < nt; ++b) {
    const int c0 = S.t_ptr[b], c1 = S.t_ptr[b + 1], T = c1 - c0;
    hl.clear();
    for (int c = c0; c < c1; ++c)
      for (int j = 0; j < 3; ++j) {
        const int nb = cn[(size_t)j * N + c];
        if (nb >= 0 && (nb < c0 || nb >= c1)) hl.push_back(nb);
      }
    std::sort(hl.begin(), hl.end());
    hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
    if (T + (int)hl.size() > cap) { S.err = "tile construction exceeded its capacity"; return STORM_E_MESH; }
    for (int k = 0; k < (int)hl.size(); ++k) hloc[hl[k]] = T + k;
    for (int c = c0; c < c1; ++c) {
      uint64_t w = 0;
      for (int j = 0; j < 3; ++j) {
        const int nb = cn[(size_t)j * N + c];
        const uint64_t l = nb < 0 ? 0xFFFFu : (nb >= c0 && nb < c1) ? (uint64_t)(nb - c0) : (uint64_t)hloc[nb];
        w |= l << (16 * j);
        const uint32_t bits = (uint32_t)(S.cmeta[c] >> (CM_SLOT0 + CM_SLOT_BITS * j));
        const uint32_t kind = bits & CMS_KIND;
        if (kind == SK_IN_Q1 || kind == SK_IN_Q5) S.t_qmask[b] |= 1ull << ((bits >> CMS_CHAIN) & 0x3F);
      }
      S.t_cln[c] = w;
    }
    for (int c : hl) S.t_hcell.push_back(c);
    S.t_hptr.push_back((int)S.t_hcell.size());
  }
  // block -> tile: each of the 8 XCDs (blocks are dealt to them round-robin)
  // takes a contiguous range of tiles, for L2 reuse of the shared halos; the
  // tiles that run an inflow fixed point go first
  const int NX = 8, q = nt / NX, r = nt % NX;
  S.t_order.resize(nt);
  for (int blk = 0; blk < nt; ++blk) {
    const int x = blk % NX, k = blk / NX;
    S.t_order[blk] = x * q + (x < r ? x : r) + k;
  }
  int front = 0;
  for (int blk = 0; blk < nt; ++blk)
    if (S.t_qmask[S.t_order[blk]]) std::swap(S.t_order[blk], S.t_order[front++]);
  // what the kernel reads first: per block the tile header, per halo entry
  // its cell's neighbours, vertices and metadata
  S.t_hdr.assign(8 * (size_t)nt, 0);
  for (int blk = 0; blk < nt; ++blk) {
    const int b = S.t_order[blk];
    int* h = &S.t_hdr[8 * (size_t)blk];
    h[0] = S.t_ptr[b]; h[1] = S.t_ptr[b + 1] - S.t_ptr[b];
    h[2] = S.t_hptr[b]; h[3] = S.t_hptr[b + 1] - S.t_hptr[b];
    h[4] = (int)(uint32_t)(S.t_qmask[b] & 0xFFFFFFFFu); h[5] = (int)(uint32_t)(S.t_qmask[b] >> 32);
    h[6] = b;
    h[7] = tile_reach[b];
  }
  // per block, the sides of its tile whose fluxes are formed once each:
  // interior edges first (owned-owned under the lower local id), then the
  // boundary sides, so that a wave mostly runs one kind
  S.t_ehdr.assign(2 * (size_t)nt, 0);
  S.t_erec.clear();
  for (int blk = 0; blk < nt; ++blk) {
    const int b = S.t_order[blk];
    const int c0 = S.t_ptr[b], T = S.t_ptr[b + 1] - c0;
    S.t_ehdr[2 * (size_t)blk] = (int)S.t_erec.size();
    int n_int = 0;
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 1) n_int = (int)S.t_erec.size() - S.t_ehdr[2 * (size_t)blk];
      for (int a = 0; a < T; ++a) {
        const int c = c0 + a;
        for (int j = 0; j < 3; ++j) {
          const uint64_t bits = (S.cmeta[c] >> (CM_SLOT0 + CM_SLOT_BITS * j)) & ((1u << CM_SLOT_BITS) - 1);
          const bool interior = (bits & CMS_KIND) == SK_INTERIOR;
          if (interior != (pass == 0)) continue;
          uint64_t k = 0;
          if (interior) {
            k = (S.t_cln[c] >> (16 * j)) & 0xFFFFu;
            if (k < (uint64_t)T && k < (uint64_t)a) continue;  // listed under the lower id
          }
          S.t_erec.push_back((uint64_t)a | ((uint64_t)j << 16) | (bits << 18) | (k << 36));
        }
      }
    }
    const int n_bnd = (int)S.t_erec.size() - S.t_ehdr[2 * (size_t)blk] - n_int;
    S.t_ehdr[2 * (size_t)blk + 1] = n_int | (n_bnd << 16);  // both below 3 * cap
  }
  const size_t nh = S.t_hcell.size();
  S.t_hrec.assign(8 * nh, 0);
  S.t_hmeta.assign(nh, 0);
  for (size_t k = 0; k < nh; ++k) {
    const int c = S.t_hcell[k];
    int* r = &S.t_hrec[8 * k];
    r[0] = c;
    for (int j = 0; j < 3; ++j) {
      r[1 + j] = cn[(size_t)j * N + c];
      r[4 + j] = S.cvtx[(size_t)j * N + c];
    }
    S.t_hmeta[k] = S.cmeta[c];
  }
  return STORM_OK;
}

}  // namespace storm
