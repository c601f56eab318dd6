// Host-side set-up of the FVT step: the mesh tables of the reference's
// find_edges/t_areas/walls/ibc_arrays/node_db/elemnt_db and the
// coefficients of dataprepFVT, laid out for the device kernels (SoA, cells in
// Reverse Cuthill-McKee order, per-cell edge slots in local order).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/storm_fvt.h"

namespace storm {

// Cell-edge slot kinds (cinfo bits 0-3).  The boundary kinds name the
// reference routine that owns the edge (invisc_fluxes.f90:195-238).
enum SlotKind : uint32_t {
  SK_INTERIOR = 0,
  SK_WALL = 1,      // wenslp3_fluxes / weslp2_fluxes
  SK_IN_Q1 = 2,     // inflowQbyH, vtype 1
  SK_IN_Q5 = 3,     // inflowQbyH, vtype 5
  SK_IN_H = 4,      // inflowH, vtype 4
  SK_OUT_H = 5,     // outflowbyH, htype 1
  SK_OUT_CRIT = 6,  // crit_outflow, htype 2
  SK_OUT_FREE = 7,  // free_flow, htype 0
  SK_NONE = 8       // boundary edge with no flux routine (inflow strings of other vtypes)
};

// cinfo layout
constexpr uint32_t CI_KIND_MASK = 0xF;
constexpr uint32_t CI_NBRLOC_SHIFT = 4;   // 2 bits: neighbour's local index of the edge
constexpr uint32_t CI_LEFT = 1u << 6;     // this cell is the edge's e(1)
constexpr uint32_t CI_HBAR = 1u << 7;     // edge of a stage (outflow) string
constexpr uint32_t CI_CHAIN_SHIFT = 8;    // 8 bits: boundary string index

constexpr int MAX_STRINGS = 64;

struct Prep {
  int N = 0, P = 0, E = 0, R = 2;
  storm_fvt_opts_t o{};
  std::vector<int> perm, iperm;     // cells: internal -> input, input -> internal
  std::vector<int> vperm, viperm;   // vertices
  // per cell (internal order); [3][N] arrays are slot-major: a[j*N + i]
  std::vector<int> cvtx, cnbr, ep1, ep2;
  std::vector<uint32_t> cinfo;
  std::vector<uint8_t> corder;      // processing order of the 3 slots, 2 bits each
  std::vector<double> rc;           // [6][N]: (rcx_j, rcy_j) rows for j = 0..2
  std::vector<double> enx, eny, etx, ety, elen, ezmid;  // [3][N]
  std::vector<double> area, perim, z, cd, delzx, delzy, zhigh;
  std::vector<double> vdx, vdy, zv3;  // [3][N] vertex offsets from the centroid, vertex beds
  std::vector<int> mptr, mcell;       // least-squares molecule (CSR)
  std::vector<double> mdx, mdy, mw, ai, bi, ci, det;
  std::vector<int> gconn;             // [3][N] Green-Gauss neighbours
  std::vector<double> gcx, gcy, gwt;  // [3][N]
  // vertices (internal order)
  std::vector<int> vptr, vcell;       // vertex -> cells (input cell order kept)
  std::vector<double> vw;             // inverse-distance weights (c2v_weights_id)
  std::vector<double> zvtx, xv_int, yv_int;  // bed and coordinates per vertex
  std::vector<int> vbc;               // stage string index of htype-1 string nodes, else -1
  // boundary strings
  int n_in = 0, n_out = 0;
  std::vector<int> in_type, out_type;
  std::vector<int> qts_ptr, hts_ptr;
  std::vector<double> qts_t, qts_v, hts_t, hts_v;
  // inflowQbyH strings (vtype 1 and 5): edge list per string in string order
  std::vector<int> qs_ptr, qs_cell, qs_slot;  // qs_ptr over n_in (empty range = other vtype)
  double bch_max = 0.0;
  int last_cell = 0;                  // internal id of input cell N-1
  // initial state (internal order)
  std::vector<double> h0, u0, v0;
  // input-order copies needed for output (input vertex ids of cells)
  std::string err;
};

// Builds everything from the C-ABI mesh; returns STORM_OK or a STORM_E_* code
// with P.err set.
int build_prep(const storm_mesh_t* m, const storm_fvt_opts_t* o, Prep& P);

// Least-squares molecule only (elemnt_db.f90:42-128), input numbering.
int build_molecule(const storm_mesh_t* m, std::vector<int>& ptr, std::vector<int>& cells,
                   std::string& err);

// Reference edge table in first-appearance order (find_edges.f90:47-54,174-187).
void build_edges(int N, int P, const int32_t* tri, std::vector<int>& ep, std::vector<int>& ee);

}  // namespace storm
