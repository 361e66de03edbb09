// Device-side data structures shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tg {

// One uniform block of cells (one ensemble member).  Cells of all instances
// are concatenated in the global SoA arrays; instance b owns
// [off, off + nx*ny).
struct InstDev {
  int nx, ny;
  long long off;
  // zero-mark blocks of the member: nsx x nsy blocks of MARK_W x MARK_H
  // cells, words [soff, soff + nsx * nsy) of a mark array
  int nsx, nsy;
  long long soff;
  double dx, dy;
  // 1/dx, 1/(dx+dx), 1/dy, 1/(dy+dy): the dual slopes' reciprocals
  // (dualmesh.C:1459), evaluated once on the host (same IEEE result)
  double inv_dx, inv_sx, inv_dy, inv_sy;
  double sin_int, tan_bed, kc1, kroot, dkroot, eps, gscale;
  double maxtime, nextoutput, nextsave;
  // first entry of the member's materials in the material table
  int mat_base;
};

// Bed friction of one material (MatProps::bedfrict[material],
// get_coef_and_eigen.C:112-114, residual.C:213): tan(phi_bed), kact's root
// sqrt|1 - (1 + tan^2 phi_bed) cos^2 phi_int| and its phi_bed derivative.
struct MatDev {
  double tan_bed, kroot, dkroot, pad;
};
constexpr int MAX_MAT = 16;  // materials per member (TG_MAX_MATERIALS)

// The bed friction of cell c of member I: its material's entry when the
// mesh gives per-cell materials (mat non-null), else the member's own.
__device__ __forceinline__ MatDev mat_of(const InstDev& I, const uint8_t* mat, const MatDev* tab,
                                         long long c) {
  if (!mat) return MatDev{I.tan_bed, I.kroot, I.dkroot, 0.};
  return tab[I.mat_base + mat[c]];
}

// A launch tile: TX x (BLOCK / TX) cells of instance `inst` starting at
// (x0, y0).  One workgroup per tile.  mw: the zero-mark word of the tile's
// block; mwW, mwE, mwS, mwN: those of the four side blocks (-1 outside the
// member).
struct TileDev {
  int inst, x0, y0, mw;
  int mwW, mwE, mwS, mwN;
};

// A run of at most BOX_CHUNK listed box cells [begin, end) of member inst
// (MaxH functional record).
struct BoxChunk {
  int inst, begin, end, pad;
};
constexpr int BOX_CHUNK = 1024;

// A mark block of one tile shape: its zero-mark words (as in TileDev) and
// the indices in that shape's tile table of its four tiles, by byte of the
// mark word (-1 where the member edge leaves no tile).  The coefficient and
// classifying passes run one workgroup per block; a shape's block table is
// indexed by mark word (entry b has mw == b).
struct BlockDev {
  int inst, mw, mwW, mwE, mwS, mwN;
  int tile[4];
  int x0, y0;  // the block's first cell in the member (multiples of MARK_W, MARK_H)
  int nx, ny;  // the member's extent and first cell (InstDev nx, ny, off), so
  long long off;  // the block passes address cells without the member record
};

// Per-instance CFL reduction cells: bit patterns of non-negative doubles
// (IEEE order of non-negative doubles equals unsigned-integer order, so
// integer atomic min/max is an exact FP64 min/max).
struct RedDev {
  unsigned long long min_ev;  // min over wet cells of min(dx,dy)/evalue
  unsigned long long hmax;    // max over wet cells of h
};

constexpr int BLOCK = 256;
// Tile variants (one tile table each): tile width TX (height BLOCK / TX)
// and whether the table is banded by XCD -- workgroup b runs on XCD b % 8,
// so a banded table hands XCD k a contiguous run of tiles (rows of the
// block), keeping halo rows in one L2.  Variant 0 is the plain row-major
// order.  The autotuner picks a variant per kernel.
constexpr int NSHAPE = 4;
constexpr int SHAPE_TX[NSHAPE] = {64, 64, 32, 16};
constexpr bool SHAPE_XCD[NSHAPE] = {false, true, true, true};
constexpr int NXCD = 8;
constexpr int DEFAULT_VARIANT = 0;

// Zero marks.  A state (h, hu, hv) or adjoint buffer can carry one 32-bit
// word per MARK_W x MARK_H block of a member's cells.  Every tile shape cuts
// a block into four tiles; byte b of the word is 1 when the b-th tile held
// +0.0 in all three fields the last time it was written, and 0 when it was
// written with anything else (each tile stores its own byte: no atomics).
// A word equal to MARK_FULL says the whole block is +0.0 (a block cut off
// by the member's edge, with fewer than four tiles of some shape, may never
// reach MARK_FULL: that only costs a skip).  Every kernel that writes a
// marked buffer keeps its words; a buffer written any other way has its
// words cleared (nothing known).
constexpr int MARK_W = 64, MARK_H = 16;
constexpr uint32_t MARK_FULL = 0x01010101u;
static_assert(MARK_W % 64 == 0 && MARK_H % (BLOCK / 16) == 0 && MARK_W * MARK_H == 4 * BLOCK,
              "every tile shape must cut a mark block into four tiles");
// Position p of n, walked by workgroup p % gridDim.x (on XCD p % NXCD when
// gridDim.x is a multiple of NXCD), reads list entry start(x) + p / NXCD of
// band x = p % NXCD, band x a contiguous run of q or q + 1 entries (q =
// n / NXCD, the first n % NXCD bands one longer): a bijection on [0, n)
// that keeps neighbouring list entries -- a block's tiles, appended
// together -- on one XCD, whose L2 then serves their shared halos.
__device__ __forceinline__ int list_band(int p, int n) {
  const int q = n / NXCD, r = n % NXCD, x = p % NXCD;
  return x * q + (x < r ? x : r) + p / NXCD;
}
// the byte of its block's word that a tile of width TX owns
template <int TX>
__device__ __forceinline__ int mark_byte(const TileDev& t) {
  constexpr int TY = BLOCK / TX;
  return ((t.y0 % MARK_H) / TY) * (MARK_W / TX) + (t.x0 % MARK_W) / TX;
}

struct Fields {
  const double* h;
  const double* hu;
  const double* hv;
};
struct Statics {
  const double* gx;
  const double* gy;
  const double* gz;
  const double* dgx;
  const double* dgy;
  const double* kx;
  const double* ky;
};

struct StepArgs {
  const InstDev* inst;
  const TileDev* tiles;
  int ntiles;
  int shape;             // tile variant of `tiles`
  const BlockDev* blocks;  // mark blocks of that variant
  int nblocks;
  Fields U;
  Statics S;
  double* kact;          // [n] forward kact of U
  RedDev* red;           // [ninst]
  double* dt_rec;        // [ninst * max_steps]
  double* time;          // [ninst]
  int max_steps;
  int step;              // step index s (U = U^s)
  double ft;             // frict_tiny of step s
  double tiny, gamma, cfl_inner, cfl_outer;
  double* h_out;
  double* hu_out;
  double* hv_out;
  int* err;              // sticky error flag (non-finite state)
  // null, or a counter the launch adds its full-path cells to: cells the
  // coefficient pass evaluates above tiny, cells of update tiles that are
  // not entirely dry
  unsigned long long* active;
  // compacted list of the tiles that take the update's full path (written
  // by the classifying pass, walked by the update) and its length
  int* tile_list;
  int* tile_count;
  // 1: k_dt clears tile_count (the update pass then launches no fill)
  int dt_clears_count;
  // zero marks of U and of the output state (both null: no marks); a tile
  // the marks show to need nothing is skipped by the coefficient and the
  // classifying pass
  const uint32_t* mk_in;
  uint32_t* mk_out;
  // null, or a counter the launch adds the cells it writes on a short path to
  unsigned long long* short_cells;
  // per-cell material index and the material table (mat null: one material
  // per member, the InstDev fields)
  const uint8_t* mat;
  const MatDev* mtab;
  // 1 when the mark bytes of U and of the output follow this launch's tile
  // shape (byte b = tile b of the block), so single tiles may be skipped;
  // 0: only whole mark words are meaningful
  int zbytes;
};

struct AdjArgs {
  const InstDev* inst;
  const TileDev* tiles;
  int ntiles;
  int shape;             // tile variant of `tiles`
  const BlockDev* blocks;  // mark blocks of that variant
  int nblocks;
  Fields U;              // U^s
  Statics S;
  Fields lam_next;       // lambda^{s+1}
  const double* dt_rec;
  int max_steps;
  int step;              // s
  double ft;
  double tiny, gamma;
  double* lam_h;         // lambda^s out
  double* lam_hu;
  double* lam_hv;
  double* dphi_part;     // [ntiles] running sum of lambda^{s+1}.dU/dphi, or null
  unsigned long long* active;  // null, or counter of cells in non-void tiles
  int* tile_list;        // compacted non-void tiles (classifying pass) and count
  int* tile_count;
  // zero marks of U^s, lambda^{s+1} and lambda^s (all null: no marks) and
  // the short-path cell counter, as in StepArgs
  const uint32_t* mk_U;
  const uint32_t* mk_lnext;
  uint32_t* mk_out;
  unsigned long long* short_cells;
  const uint8_t* mat;     // as in StepArgs
  const MatDev* mtab;
  // as StepArgs::zbytes, for the marks of U^s (u_bytes) and of the adjoint
  // buffers lambda^{s+1}, lambda^s (l_bytes)
  int u_bytes, l_bytes;
};

// Mixed-precision counterparts: FP32 arithmetic, FP32 static coefficients
// and kact; the state and adjoint are stored as S (double: FP32 increments
// applied to an FP64 state, float: FP32 storage).
template <class S>
struct FieldsT {
  const S* h;
  const S* hu;
  const S* hv;
};
struct Statics32 {
  const float* gx;
  const float* gy;
  const float* gz;
  const float* dgx;
  const float* dgy;
  const float* kx;
  const float* ky;
};
template <class S>
struct StepArgsM {
  const InstDev* inst;
  const TileDev* tiles;
  int ntiles;
  int shape;             // tile variant of `tiles`
  const BlockDev* blocks;  // (not used by these kernels)
  int nblocks;
  FieldsT<S> U;
  Statics32 S32;
  float* kact;
  RedDev* red;
  double* dt_rec;
  double* time;
  int max_steps;
  int step;
  double ft;
  double tiny, gamma;  // thresholds kept in FP64: the switches compare in W
  S* h_out;
  S* hu_out;
  S* hv_out;
  int* err;
  unsigned long long* active;  // as in StepArgs (not counted by these kernels)
  const uint8_t* mat;          // as in StepArgs
  const MatDev* mtab;
  // zero marks of U and of the output state and their byte flag, as in
  // StepArgs (null: no marks); the tiles read and store their own bytes
  const uint32_t* mk_in;
  uint32_t* mk_out;
  int zbytes;
};
template <class S>
struct AdjArgsM {
  const InstDev* inst;
  const TileDev* tiles;
  int ntiles;
  int shape;             // tile variant of `tiles`
  const BlockDev* blocks;  // (not used by these kernels)
  int nblocks;
  FieldsT<S> U;
  Statics32 S32;
  FieldsT<S> lam_next;
  const double* dt_rec;
  int max_steps;
  int step;
  double ft;
  double tiny, gamma;
  S* lam_h;
  S* lam_hu;
  S* lam_hv;
  double* dphi_part;
  int2* thin_list;       // [ncell] (tile, thread) of cells gathered in FP64
  int* thin_count;
  const double* s64;     // FP64 static coefficients [7 * ncell] or null
  long long ncell;
  unsigned long long* active;  // null, or counter of cells in non-void tiles
  int* tile_list;        // compacted non-void tiles (classifying pass) and count
  int* tile_count;
  const uint8_t* mat;    // as in StepArgs
  const MatDev* mtab;
  // zero marks and their byte flags, as in AdjArgs (all null: no marks)
  const uint32_t* mk_U;
  const uint32_t* mk_lnext;
  uint32_t* mk_out;
  int u_bytes, l_bytes;
};
// workgroups of the FP64 thin-layer adjoint pass (strides over its list)
constexpr int THIN_GRID = 1024;
// workgroups of the update / adjoint passes that stride over the compacted
// list of tiles taking the full path (a fixed grid, for graph capture):
// 16 per CU, so that the workgroups left without a tile exit early while
// the listed tiles spread over every CU (1024, 2048, 4096 and 8192 measured
// on the 256-member ensemble: 4096 about 1 % ahead)
constexpr int LIST_GRID = 4096;

template <class S>
void launch_coef_m(const StepArgsM<S>& a, bool reduce, hipStream_t s);
template <class S>
void launch_update_m(const StepArgsM<S>& a, hipStream_t s);
template <class S>
void launch_adjoint_m(const AdjArgsM<S>& a, hipStream_t s);
void launch_stats32(const InstDev* inst, const TileDev* tiles, int ntiles, FieldsT<float> U,
                    double* part_mass, double* part_hmax, double tiny, hipStream_t s);

// Adaptive mesh (flattened leaf tables), one thread per leaf.  Sides 0 S,
// 1 E, 2 N, 3 W; slot k < 4 the neighbour on side k, slot k + 4 the second
// neighbour of side k.
struct LeafDev {
  int n;
  const int* nbr;        // [8n] slot-major: nbr[k * n + c]
  // per leaf: bits 2s..2s+1 the kind of side s (LEAF_*); bits 8+3k..10+3k
  // the slot of this leaf in the table of its slot-k neighbour
  // (Element::which_neighbor, element2.C:574-583)
  const uint32_t* meta;
  const double* cdx;     // [2n] cell size: dx at [c], dy at [n + c]
  const double* xc;      // [2n] centre: x at [c], y at [n + c]
  // zero marks: one word per chunk of BLOCK leaves (the kernels' workgroup),
  // MARK_FULL when the chunk held +0.0 in all three fields when last
  // written; the chunks holding neighbours of chunk q's leaves (q itself
  // excluded) are cnb[cnb_off[q] .. cnb_off[q + 1])
  const int* cnb_off;
  const int* cnb;
};
constexpr uint32_t LEAF_BOUNDARY = 0;  // domain boundary (neigh_proc INIT)
constexpr uint32_t LEAF_SAME = 1;      // one neighbour of the same generation
constexpr uint32_t LEAF_COARSER = 2;   // one coarser neighbour (node S_S_CON)
constexpr uint32_t LEAF_FINER = 3;     // two finer neighbours (node S_C_CON)

// Reverse-step buffers of the leaf path.
struct LeafBufs {
  double* kd;    // [n] kact from the dual slopes
  double* dkd;   // [n] d kact / d phi_bed
  // [81n] gather terms: record slot m of leaf c at jac + (m * n + c) * 9 holds
  // lambda[l] * J[l][k] of the block leaf c adds in its slot m
  double* jac;
  // per chunk: bit 0 set when kd holds getcoef's dry value (epsilon) for
  // every leaf of the chunk, bit 1 when dkd holds 0 for every leaf
  uint32_t* kdmk;
};

// Error indicator on the error grid (tg_error.hip): interpolation rules of
// tg_host_error_stencil, and the buffers of one evaluation.
constexpr int ERR_COPY = 0, ERR_LINEAR_X = 1, ERR_LINEAR_Y = 2, ERR_BARYCENTRIC = 3,
              ERR_BILINEAR = 4;
struct ErrArgs {
  int n, nfwd;           // error elements, leaves of the forward mesh
  LeafDev L;             // error grid tables
  Statics S;             // error grid static coefficients
  InstDev I;             // kact / friction constants (member 0)
  const uint8_t* mat;    // error elements' materials (null: one material)
  const MatDev* mtab;
  const int* father;     // [n] forward leaf of each element
  const int* kind;       // [n] interpolation rule
  const int* pts;        // [4n] forward leaves of the stencil points
  const double* pxy;     // [8n] their coordinates
  const double* fwd;     // [9 nfwd] previous state, state, adjoint of the forward leaves
  double* rec;           // [9n] reconstructed previous state, state, adjoint
  double* kact;          // [n]
  double* res;           // [3n] residual
  double* err;           // [n] error (el_error[1])
  double* corr;          // [n] correction
  double dt, ft, tiny, gamma;
};
// the three kernels of one evaluation; ev (4 events, or null) bracket them
void launch_error(const ErrArgs& a, hipStream_t s, hipEvent_t* ev);

void launch_leaf_coef(const StepArgs& a, const LeafDev& L, bool reduce, hipStream_t s);
void launch_leaf_update(const StepArgs& a, const LeafDev& L, hipStream_t s);
// one reverse step s: the dual kact of U^s (unless a previous launch made
// it: dualk false), the Jacobian products, the gradient term, the gather;
// with next (the arguments of step s - 1), the gather launch also makes the
// dual kact of U^{s-1}
void launch_leaf_adjoint(const AdjArgs& a, const AdjArgs* next, bool dualk, const LeafDev& L,
                         const LeafBufs& B, hipStream_t s);
void launch_leaf_stats(const LeafDev& L, Fields U, double* part_mass, double* part_hmax,
                       double tiny, hipStream_t s);

void launch_coef(const StepArgs& a, bool reduce, hipStream_t s);
void launch_dt(const StepArgs& a, int ninst, hipStream_t s);
void launch_update(const StepArgs& a, hipStream_t s);
void launch_adjoint(const AdjArgs& a, hipStream_t s);
void launch_reset_red(RedDev* red, int ninst, hipStream_t s);
// rec[member * max_steps + step] = max(itself, largest h > 0 over the listed
// box cells); rec holds FP64 values >= +0.0
template <class S>
void launch_box_max(const BoxChunk* ch, int nchunks, const long long* cells, const S* h,
                    double* rec, int max_steps, int step, hipStream_t s);
// copy (src) or zero (src null) one member's cells [off, off + count) of the
// three fields of a state / adjoint buffer; a copy clears the member's mark
// words [mw0, mw0 + nmw) of mk
void launch_member_fill(const void* src, void* dst, int es, long long ncell, long long off,
                        long long count, uint32_t* mk, long long mw0, long long nmw,
                        hipStream_t s);
// bytes must be a multiple of 4 (fields are FP32 or FP64)
void launch_zero(void* p, size_t bytes, hipStream_t s);
void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// per-tile mass and largest h of variant-0 tiles; blocks whose mark word in
// mk (null: none) is MARK_FULL are not read
void launch_stats(const InstDev* inst, const TileDev* tiles, int ntiles, Fields U,
                  const uint32_t* mk, double* part_mass, double* part_hmax, double tiny,
                  hipStream_t s);
// copy of a state buffer's three fields (src field pointers, dst the copy's
// base, fields ncell elements of es bytes apart) by the mark blocks of tile
// variant `shape` (one per mark word), writing both buffers' zero marks from
// the copied values (dst and mk_dst null: the source's marks only)
void launch_copy_marking(const BlockDev* blocks, int nblocks, int shape, int es,
                         const void* const src[3], void* dst, long long ncell, uint32_t* mk_src,
                         uint32_t* mk_dst, hipStream_t s);
// per-member reduction of per-tile partials (member b's tiles: idx[off[b] ..
// off[b+1])): out[b] sums, out[n + b] maxima (max_part given), out[2n + b] =
// time[b] (given), out[3n + b] = dt_rec[b * stride + col] (given)
void launch_member_reduce(const double* sum_part, const double* max_part, const int* off,
                          const int* idx, int n, const double* time, const double* dt_rec,
                          long long stride, int col, double* out, hipStream_t s);

}  // namespace tg
