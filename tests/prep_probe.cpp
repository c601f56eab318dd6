// Host-only probe of the set-up tables (no device code): builds the prep and
// stage tiles of a mesh and checks the tile invariants the stage kernel
// relies on.  Compiled by tests/test_prep_tiles.py with g++.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../storm_amd/csrc/fvt_prep.h"

using namespace storm;

extern "C" int probe_tiles(const storm_mesh_t* m, const storm_fvt_opts_t* o, int cap,
                           int* stats, char* msg, int msglen) {
  Prep P;
  auto fail = [&](const std::string& s) {
    std::snprintf(msg, msglen, "%s", s.c_str());
    return -100;
  };
  int rc = build_prep(m, o, P);
  if (rc != STORM_OK) { std::snprintf(msg, msglen, "%s", P.err.c_str()); return rc; }
  rc = build_tiles(P, cap);
  if (rc != STORM_OK) { std::snprintf(msg, msglen, "%s", P.err.c_str()); return rc; }
  const int N = P.N, nt = P.nt;
  if ((int)P.t_ptr.size() != nt + 1 || P.t_ptr[0] != 0 || P.t_ptr[nt] != N)
    return fail("tiles do not partition the cells");
  std::vector<int> seen(nt, 0);
  for (int b = 0; b < nt; ++b) {
    const int t = P.t_order[b];
    if (t < 0 || t >= nt || seen[t]++) return fail("block order is not a permutation");
  }
  int maxw = 0;
  long halo = 0;
  for (int b = 0; b < nt; ++b) {
    const int c0 = P.t_ptr[b], T = P.t_ptr[b + 1] - c0;
    const int hb = P.t_hptr[b], H = P.t_hptr[b + 1] - hb;
    if (T <= 0) return fail("empty tile");
    if (T + H > cap) return fail("tile over capacity");
    maxw = std::max(maxw, T + H);
    halo += H;
    uint64_t qm = 0;
    for (int c = c0; c < c0 + T; ++c) {
      for (int j = 0; j < 3; ++j) {
        const int nb = P.cnbr[(size_t)j * N + c];
        const int k = (int)(P.t_cln[c] >> (16 * j)) & 0xFFFF;
        if (nb < 0) {
          if (k != 0xFFFF) return fail("boundary side with a local neighbour");
        } else if (k < T) {
          if (c0 + k != nb) return fail("bad local id of a tile neighbour");
        } else if (k - T >= H || P.t_hcell[hb + k - T] != nb) {
          return fail("bad local id of a halo neighbour");
        }
        const uint32_t bits = (uint32_t)(P.cmeta[c] >> (CM_SLOT0 + CM_SLOT_BITS * j));
        const uint32_t kind = bits & CMS_KIND;
        if (kind == SK_IN_Q1 || kind == SK_IN_Q5) qm |= 1ull << ((bits >> CMS_CHAIN) & 0x3F);
      }
    }
    if (qm != P.t_qmask[b]) return fail("inflow string mask mismatch");
    for (int k = 0; k < H; ++k) {
      const int c = P.t_hcell[hb + k];
      if (c >= c0 && c < c0 + T) return fail("halo cell inside the tile");
    }
  }
  for (int blk = 0; blk < nt; ++blk) {
    const int b = P.t_order[blk];
    const int* h = &P.t_hdr[8 * (size_t)blk];
    const uint64_t qm = (uint64_t)(uint32_t)h[4] | ((uint64_t)(uint32_t)h[5] << 32);
    if (h[0] != P.t_ptr[b] || h[1] != P.t_ptr[b + 1] - P.t_ptr[b] || h[2] != P.t_hptr[b] ||
        h[3] != P.t_hptr[b + 1] - P.t_hptr[b] || qm != P.t_qmask[b] || h[6] != b)
      return fail("bad block header");
  }
  for (size_t k = 0; k < P.t_hcell.size(); ++k) {
    const int c = P.t_hcell[k];
    const int* r = &P.t_hrec[8 * k];
    bool ok = r[0] == c && P.t_hmeta[k] == P.cmeta[c];
    for (int j = 0; j < 3; ++j)
      ok = ok && r[1 + j] == P.cnbr[(size_t)j * N + c] && r[4 + j] == P.cvtx[(size_t)j * N + c];
    if (!ok) return fail("bad halo record");
  }
  // the per-reach lists of stage-string vertices hold exactly the vertices
  // with a stage string, once each
  if ((int)P.r_svptr.size() != P.nr + 1 || P.r_svptr[0] != 0 ||
      P.r_svptr[P.nr] != (int)P.r_svtx.size())
    return fail("bad stage-string vertex lists");
  std::vector<int> listed(P.P, 0);
  for (int v : P.r_svtx) {
    if (v < 0 || v >= P.P || P.vbc[v] < 0) return fail("listed vertex without a stage string");
    listed[v]++;
  }
  for (int v = 0; v < P.P; ++v)
    if ((P.vbc[v] >= 0) != (listed[v] == 1)) return fail("stage-string vertex not listed once");
  stats[0] = nt; stats[1] = maxw; stats[2] = (int)P.r_svtx.size(); stats[3] = (int)halo;
  return 0;
}
