// Host-side set-up of the FVT step: the mesh tables of the reference's
// find_edges/t_areas/walls/ibc_arrays/node_db/elemnt_db and the
// coefficients of dataprepFVT, laid out for the device kernels (cells in
// Hilbert-curve order, per-cell edge slots in local order), and the tiles of
// the stage kernel.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/storm_fvt.h"

namespace storm {

// Cell-edge slot kinds.  The boundary kinds name the reference routine that
// owns the edge (invisc_fluxes.f90:195-238).
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

// Per-cell metadata word (uint64):
//   bits 0-5   processing order of the three local edges, 2 bits each: the
//              reference's phi accumulation order (interior edges by edge
//              number, then walls, then the boundary routines in call order)
//   slot j at bit 6 + 18*j (j = 0..2, local edge j runs from vertex j to j+1):
//     +0  kind (4 bits)        +4  neighbour's local index of the edge (2)
//     +6  cell is e(1) (1)     +7  edge of a stage string, hvalve (1)
//     +8  edges%p runs from vertex j+1 to j (1)   +9 boundary string (8)
constexpr int CM_SLOT0 = 6, CM_SLOT_BITS = 18;
constexpr uint64_t CMS_KIND = 0xF;
constexpr int CMS_NL = 4, CMS_LEFT = 6, CMS_HBAR = 7, CMS_REV = 8, CMS_CHAIN = 9;
// bit 60: the cell is the inlet or outlet of a culvert
constexpr int CM_CULVERT = 60;

constexpr int MAX_STRINGS = 64;  // boundary strings of each kind per reach
// Most cells a tile of capacity cap (cells + halo) owns.  The stage kernel
// sizes the owned cells' vertex arrays in shared memory by it.  At capacity
// 256 it is 224, which keeps a block's shared memory under 40 KB, so four
// blocks (four waves per SIMD) share a CU; a tile rarely owns more, since its
// halo is about a quarter of its owned cells.  The larger capacities run one
// or two blocks per CU and keep the whole capacity.
constexpr int tile_tmax(int cap) { return cap == 256 ? 224 : cap; }

// A batch of independent reaches laid end to end in one mesh: reach r owns
// input cells [cell_ptr[r], cell_ptr[r+1]), inflow strings [in_ptr[r],
// in_ptr[r+1]) and outflow strings [out_ptr[r], out_ptr[r+1]).  No edge joins
// two reaches.
struct ReachSplit {
  std::vector<int> cell_ptr, in_ptr, out_ptr;
};

struct Prep {
  int N = 0, P = 0, E = 0, R = 2;
  storm_fvt_opts_t o{};
  std::vector<int> perm, iperm;     // cells: internal -> input, input -> internal
  std::vector<int> vperm, viperm;   // vertices
  // per cell (internal order); [3][N] arrays are slot-major: a[j*N + i]
  std::vector<int> cvtx, cnbr;
  std::vector<uint64_t> cmeta;
  std::vector<double> xc, yc, z, cd;
  std::vector<int> mptr, mcell;       // least-squares molecule (CSR)
  std::vector<int> gconn;             // [3][N] Green-Gauss neighbours
  std::vector<double> gcx, gcy, gwt;  // [3][N]
  // vertices (internal order)
  std::vector<double> vxyz;           // [4][P] interleaved: x, y, zvtx, 0
  std::vector<int> vptr, vcell;       // vertex -> cells (input cell order kept)
  std::vector<int> vbc;               // stage string index of htype-1 string nodes, else -1
  int vW = 0;                         // largest vertex degree
  std::vector<int> vell;              // [vW][P] incident cells in n2t order, -1 past the degree
  std::vector<double> vwt;            // [vW][P] their inverse-distance weights
  std::vector<double> vzb;            // [2][P] interleaved: zvtx, bits (vbc | degree << 32 | reach << 40)
  // boundary strings
  int n_in = 0, n_out = 0;
  std::vector<int> in_type, out_type;
  std::vector<int> qts_ptr, hts_ptr;
  std::vector<double> qts_t, qts_v, hts_t, hts_v;
  // inflowQbyH strings (vtype 1 and 5): edges in string order with their cell,
  // local edge and geometry
  std::vector<int> qs_ptr, qs_cell, qs_slot;
  std::vector<double> qs_rcx, qs_rcy, qs_nx, qs_ny, qs_len;
  // reaches (one unless a batch): internal cells [r_cell[r], r_cell[r+1]),
  // first string of each kind, the internal id of the reach's last input
  // cell (its local dt drives steady friction), bch_max of dataprepFVT
  int nr = 1;
  std::vector<int> r_cell, r_in, r_out, r_last, r_tile;
  // per reach, the vertices of its stage strings (vbc >= 0):
  // r_svtx[r_svptr[r] .. r_svptr[r+1])
  std::vector<int> r_svptr, r_svtx;
  std::vector<double> r_bch;
  std::vector<double> h0, u0, v0;     // initial state (internal order)
  // The final to_vertices' velocity part (to_vertices.f90:56-65 with
  // enforce_bcs.f90).  Per vertex (internal order): the outflow-string node
  // normals of ibc_arrays.f90:183-253 its velocity is checked against
  // (htype 0, or htype 1 with hvalve) as vo_nrm[2*vo_ptr[p] ..
  // 2*vo_ptr[p+1]), and the index of its wall tangent (ibc_arrays.f90:
  // 266-330) in wtang, -1 if it is no wall point (walls.f90:150-176).
  std::vector<int> vo_ptr, vwall;
  std::vector<double> vo_nrm, wtang;
  // Every inflow string's nodes (internal ids) in string order, with the
  // node normals of ibc_arrays.f90:93-135 (interleaved x, y) and, at each
  // node but the last, the length of the edge to the next node (l_inflow).
  std::vector<int> is_ptr, is_node;
  std::vector<double> is_pn, is_len;
  // dataprepFVT.f90:73-140's inputs given at vertices, in input order (empty
  // when given at cells), for the module copies it leaves
  std::vector<double> vin_h, vin_u, vin_v, vin_cd;
  // node_db.f90:165-185: n2n, the vertices sharing an edge with vertex p
  // (internal ids, in edge-table order) as nn_idx[nn_ptr[p] .. nn_ptr[p+1]),
  // for the node routines of the RDS solver (storm_fvt_fvsources)
  std::vector<int> nn_ptr, nn_idx;
  // stage-kernel tiles: tile b owns cells [t_ptr[b], t_ptr[b+1]); its halo
  // (edge neighbours outside) is listed per tile; local cell ids are 16
  // bits, three to a word (cell slot j at bit 16*j): tile cells 0..T-1, then
  // halo cells in list order
  int cap = 0, nt = 0;
  std::vector<int> t_order;           // block -> tile
  std::vector<int> t_ptr, t_hptr, t_hcell;
  std::vector<uint64_t> t_cln;        // per cell: local neighbour ids (0xFFFF: none)
  std::vector<uint64_t> t_qmask;      // per tile: inflowQbyH strings it touches
  // device copies in block order: 8 ints per block {c0, T, hb, H, qmask lo,
  // qmask hi, tile, 0}; 8 ints per halo entry {cell, nbr0..2, vtx0..2, 0}
  // and its metadata word
  std::vector<int> t_hdr, t_hrec;
  std::vector<uint64_t> t_hmeta;
  // Per block, the sides whose fluxes the stage kernel forms once each
  // (2 ints per block {first record, interior count | boundary count << 16};
  // one uint64 per record): every
  // interior edge of the tile with at least one owned cell, then every
  // boundary side of an owned cell.  Record: bits 0-15 owned local id a,
  // 16-17 its side ja, 18-35 the slot bits of (a, ja), 36-51 the local id of
  // the neighbour across an interior side (an interior edge between two owned
  // cells appears once, under the lower id).
  std::vector<int> t_ehdr;
  std::vector<uint64_t> t_erec;
  // Residual smoothing (opts rsavg): the Gauss-Seidel sweep of
  // solverFVT.f90:244-255 as levels of independent cells.  Level l holds
  // rs_cell[rs_lptr[l] .. rs_lptr[l+1]) (internal ids, ascending); reach r
  // owns levels [r_lev[r], r_lev[r+1]).  Position k's t2t row (elemnt_db
  // order) is rs_nbr[rs_nptr[k] .. rs_nptr[k+1]), internal ids with bit 31
  // set where the neighbour comes later in input order (its old value).
  std::vector<int> rs_lptr, rs_cell, rs_nptr, rs_nbr, r_lev;
  // The sweep through shared memory (rs_fast; when the mesh fits its
  // limits): every value a level reads lives in a block-local window, so a
  // level waits on no global memory.  Position p's value sits in ring slot
  // (p - first position of its reach) % RSF_W from RSF_G + 1 levels before
  // p's own level (filled with the previous sweep's value) until the ring
  // wraps; a value read outside that window goes through one of RSF_FS far
  // slots (shared-memory index RSF_W + slot).  rs_code[e * N + k]: the
  // shared-memory index of entry e of position k's t2t row, -1 past its end;
  // rs_farw[k]: the far slot position k also writes, or -1.  Level l fills
  // the ring with level l + RSF_G + 1's positions and the far slots listed
  // in rs_fsrc/rs_fdst[rs_fptr[l] .. rs_fptr[l+1]) (source position, far
  // slot); each reach first fills rs_psrc/rs_pdst[rs_pptr[r] ..
  // rs_pptr[r+1]) (source position, shared-memory index).
  // rs_hdr: per level l two int4, {first position, width, first position of
  // level l + RSF_G + 1 (the ring fill), its width or 0 past the reach},
  // {rs_fptr[l], its count, 0, 0}; then one zero entry (index L).
  bool rs_fast = false;
  int rs_nfar = 0;
  std::vector<int> rs_code, rs_farw, rs_fptr, rs_fsrc, rs_fdst, rs_pptr, rs_psrc, rs_pdst, rs_hdr;
  std::string err;
};

// Limits of the shared-memory sweep: threads (positions per level), ring
// slots, lookahead levels of the ring fill, far slots, entries per row.
constexpr int RSF_NT = 512, RSF_W = 4096, RSF_G = 2, RSF_FS = 1536, RSF_E = 8;
static_assert(RSF_W + RSF_FS < 32768, "entry codes are stored in 16 bits");

// Builds everything from the C-ABI mesh; returns STORM_OK or a STORM_E_* code
// with P.err set.
int build_prep(const storm_mesh_t* m, const storm_fvt_opts_t* o, Prep& P,
               const ReachSplit* split = nullptr);

// Tiles of at most `cap` cells + halo cells, none spanning two reaches
// (needs the cell tables of build_prep).
int build_tiles(Prep& P, int cap);

// The module copies dataprepFVT.f90:69-140 leaves, in input order (any
// pointer may be null): cells z, zeta, cd; vertices zvtx, cdvtx, hvtx, uvtx,
// vvtx, zetavtx.  Cell values given at vertices are averaged by vtx2ctr
// (vtx2ctr.f90:32-35); vertex values given at cells are interpolated by
// ctr2vtx type 3 (ctr2vtx.f90:73-80 with c2v_setup.f90:106-119).
void dataprep_fields(const Prep& P, double* z, double* zvtx, double* zeta, double* cd,
                     double* cdvtx, double* hvtx, double* uvtx, double* vvtx, double* zetavtx);

// windforcing.f90:64-86 (opt_solver 2, 3): per cell (internal order) the
// pair wcoef0 w |w| / rho (x then y), from n == N cell values (input order)
// or n == P vertex values averaged by vtx2ctr.f90:32-35.
int wind_coefs(const Prep& P, int64_t n, const double* c, const double* wx, const double* wy,
               std::vector<double>& w, std::string& err);

// cvprep.f90:29-58 with pto_in_triangle.f90:20-43: per culvert the first
// cell in input order holding the inlet (and the outlet) point, as input
// ids, and the inlet's offset from that cell's centroid (x, y).
int locate_culverts(const Prep& P, int n, const double* pin, const double* pout,
                    std::vector<int>& cin, std::vector<int>& cout, std::vector<double>& rin,
                    std::string& err);

// Least-squares molecule only (elemnt_db.f90:42-128), input numbering.
int build_molecule(const storm_mesh_t* m, std::vector<int>& ptr, std::vector<int>& cells,
                   std::string& err);

// Reference edge table in first-appearance order (find_edges.f90:47-54,174-187).
void build_edges(int N, int P, const int32_t* tri, std::vector<int>& ep, std::vector<int>& ee);

}  // namespace storm
