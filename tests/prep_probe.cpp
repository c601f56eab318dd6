// Host-only probe of the set-up tables (no device code): builds the prep and
// stage tiles of a mesh and checks the tile invariants the stage kernel
// relies on.  Compiled by tests/test_prep_tiles.py with g++.
#include <cstdio>
#include <algorithm>
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

// The residual smoothing's level schedule and the shared-memory sweep's
// tables, run on the host: nsweep sweeps of random phi0 (3 components) by
//   (1) the serial Gauss-Seidel sweep in input cell order over the t2t rows,
//   (2) levels of the global-memory kernel (rs_lptr/rs_cell/rs_nbr),
//   (3) the shared-memory kernel's steps: per level every position's value
//       from its ring or far slots, its writes, then the level's fills.
// (3) tags every slot with the position and sweep whose value it holds and
// fails on a read of a stale or foreign value, or of a slot written in the
// same level (a race between the level's threads).  stats: fast flag,
// levels, far slots, max level width; returns 0 when (1) = (2) = (3) bit
// for bit.
extern "C" int probe_rsavg(const storm_mesh_t* m, const storm_fvt_opts_t* o, int nsweep,
                           unsigned seed, int* stats, char* msg, int msglen) {
  Prep P;
  auto fail = [&](const std::string& s) {
    std::snprintf(msg, msglen, "%s", s.c_str());
    return -100;
  };
  int rc = build_prep(m, o, P);
  if (rc != STORM_OK) { std::snprintf(msg, msglen, "%s", P.err.c_str()); return rc; }
  const int N = P.N;
  const double eps = 0.5, fac1 = 1.0 / (1.0 + eps * 3.0), fac2 = eps / (1.0 + eps * 3.0);
  std::vector<double> phi0(3 * (size_t)N);
  uint64_t x = 0x9E3779B97F4A7C15ull ^ seed;
  for (auto& v : phi0) {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    v = (double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  // (1) serial, input order, rows of the molecule (t2t for LSQ gradients)
  if (o->gradient == STORM_GRAD_HDLS) return fail("probe needs the t2t molecule (LSQ)");
  std::vector<double> a(phi0);
  for (int s = 0; s < nsweep; ++s)
    for (int i = 0; i < N; ++i) {
      const int ii = P.iperm[i];
      for (int c = 0; c < 3; ++c) {
        double q = fac1 * phi0[(size_t)c * N + ii];
        for (int t = P.mptr[ii]; t < P.mptr[ii + 1]; ++t)
          q = q + fac2 * a[(size_t)c * N + P.mcell[t]];
        a[(size_t)c * N + ii] = q;
      }
    }
  // (2) levels, global memory
  const int NR = (int)P.r_lev.size() - 1, L = P.r_lev[NR];
  std::vector<double> b(phi0);
  for (int s = 0; s < nsweep; ++s)
    for (int l = 0; l < L; ++l) {
      std::vector<double> lvl;
      for (int k = P.rs_lptr[l]; k < P.rs_lptr[l + 1]; ++k) {
        const int i = P.rs_cell[k];
        for (int c = 0; c < 3; ++c) {
          double q = fac1 * phi0[(size_t)c * N + i];
          for (int m2 = P.rs_nptr[k]; m2 < P.rs_nptr[k + 1]; ++m2) {
            const int e = P.rs_nbr[m2];
            const int j = e & 0x7fffffff;
            q = q + fac2 * ((s == 0 && e < 0) ? phi0 : b)[(size_t)c * N + j];
          }
          lvl.push_back(q);
        }
      }
      size_t u = 0;  // a level's values are written after all its reads
      for (int k = P.rs_lptr[l]; k < P.rs_lptr[l + 1]; ++k)
        for (int c = 0; c < 3; ++c) b[(size_t)c * N + P.rs_cell[k]] = lvl[u++];
    }
  for (size_t q = 0; q < a.size(); ++q)
    if (a[q] != b[q]) return fail("level sweep differs from the serial sweep");
  int maxw = 0;
  for (int l = 0; l < L; ++l) maxw = std::max(maxw, P.rs_lptr[l + 1] - P.rs_lptr[l]);
  stats[0] = P.rs_fast; stats[1] = L; stats[2] = P.rs_nfar; stats[3] = maxw;
  if (!P.rs_fast) return 0;
  // (3) the shared-memory kernel
  const int W = RSF_W, G = RSF_G, S = W + RSF_FS;
  std::vector<int> pos(N);
  for (int k = 0; k < N; ++k) pos[P.rs_cell[k]] = k;
  std::vector<double> lo_prev(3 * (size_t)N), lo_next(3 * (size_t)N), outc(3 * (size_t)N);
  for (int k = 0; k < N; ++k)
    for (int c = 0; c < 3; ++c) lo_prev[(size_t)c * N + k] = phi0[(size_t)c * N + P.rs_cell[k]];
  const std::vector<double> p0lo(lo_prev);
  for (int s = 0; s < nsweep; ++s) {
    for (int r = 0; r < NR; ++r) {
      const int l0 = P.r_lev[r], l1 = P.r_lev[r + 1], base = P.rs_lptr[l0];
      std::vector<double> sm(3 * (size_t)S, 0.0);
      // tag: position * 2 + (1: this sweep's value, 0: previous), -1 empty
      std::vector<long> tag(S, -1);
      auto put = [&](int slot, int p, int ver, const double* v, long vs, int) {
        for (int c = 0; c < 3; ++c) sm[3 * (size_t)slot + c] = v[c * vs];
        tag[slot] = 2L * p + ver;
      };
      for (int mm = P.rs_pptr[r]; mm < P.rs_pptr[r + 1]; ++mm)
        put(P.rs_pdst[mm], P.rs_psrc[mm], 0, &lo_prev[P.rs_psrc[mm]], N, l0 - 1);
      for (int l = l0; l < l1; ++l) {
        // step 1 (all reads of the level before any of its writes)
        std::vector<double> val;
        std::vector<int> rd;
        for (int k = P.rs_lptr[l]; k < P.rs_lptr[l + 1]; ++k) {
          const int i = P.rs_cell[k];
          double q[3];
          for (int c = 0; c < 3; ++c) q[c] = fac1 * p0lo[(size_t)c * N + k];
          int e = 0;
          for (int m2 = P.rs_nptr[k]; m2 < P.rs_nptr[k + 1]; ++m2, ++e) {
            const int code = P.rs_code[(size_t)e * N + k];
            if (code < 0 || code >= S) return fail("entry without a slot");
            const int ent = P.rs_nbr[m2];
            const int p = pos[ent & 0x7fffffff];
            const long want = 2L * p + (ent < 0 ? 0 : 1);
            if (tag[code] != want) return fail("slot holds another value than the entry's");
            rd.push_back(code);
            for (int c = 0; c < 3; ++c) q[c] = q[c] + fac2 * sm[3 * (size_t)code + c];
          }
          if (e < RSF_E && P.rs_code[(size_t)e * N + k] != -1) return fail("row not ended");
          (void)i;
          for (int c = 0; c < 3; ++c) val.push_back(q[c]);
        }
        // step 1 writes, then the fills (each slot at most once per level)
        std::vector<int> hit;
        size_t u = 0;
        for (int k = P.rs_lptr[l]; k < P.rs_lptr[l + 1]; ++k, u += 3) {
          const int slot = (k - base) % W;
          hit.push_back(slot);
          put(slot, k, 1, &val[u], 1, l);
          if (P.rs_farw[k] >= 0) { hit.push_back(W + P.rs_farw[k]); put(W + P.rs_farw[k], k, 1, &val[u], 1, l); }
          for (int c = 0; c < 3; ++c) {
            lo_next[(size_t)c * N + k] = val[u + c];
            outc[(size_t)c * N + P.rs_cell[k]] = val[u + c];
          }
        }
        if (l + G + 1 < l1)
          for (int q = P.rs_lptr[l + G + 1]; q < P.rs_lptr[l + G + 2]; ++q) {
            hit.push_back((q - base) % W);
            put((q - base) % W, q, 0, &lo_prev[q], N, l);
          }
        for (int f = P.rs_fptr[l]; f < P.rs_fptr[l + 1]; ++f) {
          hit.push_back(W + P.rs_fdst[f]);
          put(W + P.rs_fdst[f], P.rs_fsrc[f], 0, &lo_prev[P.rs_fsrc[f]], N, l);
        }
        std::sort(hit.begin(), hit.end());
        if (std::adjacent_find(hit.begin(), hit.end()) != hit.end())
          return fail("two writes to one slot in a level");
        std::sort(rd.begin(), rd.end());
        for (int sl : rd)
          if (std::binary_search(hit.begin(), hit.end(), sl))
            return fail("slot written in the level that reads it");
        if (P.rs_fptr[l + 1] - P.rs_fptr[l] > RSF_NT || P.rs_lptr[l + 1] - P.rs_lptr[l] > RSF_NT)
          return fail("a level exceeds the block");
      }
    }
    lo_prev.swap(lo_next);
  }
  for (size_t q = 0; q < a.size(); ++q)
    if (a[q] != outc[q]) return fail("shared-memory sweep differs from the serial sweep");
  return 0;
}

// The source-term set-up helpers on a mesh: wind coefficients from n values
// (cells or vertices) and culvert location; returns their status codes
// (wind in rc[0], culverts in rc[1]) and the located cells.
extern "C" int probe_sources(const storm_mesh_t* m, const storm_fvt_opts_t* o, int64_t nw,
                             const double* c, const double* wx, const double* wy, int ncv,
                             const double* pin, const double* pout, int* rc, int* cells) {
  Prep P;
  int r = build_prep(m, o, P);
  if (r != STORM_OK) return r;
  std::vector<double> w;
  std::string err;
  rc[0] = wind_coefs(P, nw, c, wx, wy, w, err);
  std::vector<int> cin, cout;
  std::vector<double> rin;
  rc[1] = ncv > 0 ? locate_culverts(P, ncv, pin, pout, cin, cout, rin, err) : 0;
  if (rc[1] == STORM_OK)
    for (int k = 0; k < ncv; ++k) { cells[2 * k] = cin[k]; cells[2 * k + 1] = cout[k]; }
  std::vector<double> z(P.N), zv(P.P), ze(P.N), cd(P.N), cdv(P.P), hv(P.P), uv(P.P), vv(P.P),
      zev(P.P);
  dataprep_fields(P, z.data(), zv.data(), ze.data(), cd.data(), cdv.data(), hv.data(), uv.data(),
                  vv.data(), zev.data());
  return 0;
}
