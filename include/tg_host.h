/*
 * tg_host.h -- host-side set-up that feeds tg_set_mesh / tg_set_state:
 * synthetic DEMs, the per-cell static coefficients the reference derives
 * from a GIS raster, and the initial pile heights.  Plain C ABI, host only.
 */
#ifndef TG_HOST_H
#define TG_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A GRASS-style elevation raster: row 0 is the northern edge (ymax), column 0
 * the western edge (xmin); values are float32 like the reference's fcell
 * reader (GisApi.C:153-285). */
typedef struct tg_dem {
  int nrows, ncols;
  double xmin, ymax, res;
  const float* elev;    /* [nrows*ncols], row-major */
} tg_dem;

/* Reads a GRASS elevation raster (the reference's GIS input, GisApi.C:153-285):
 * <gisdbase>/<location>/<mapset>/cellhd/<name> (proj, zone, north, south,
 * east, west, cols, rows, e-w resol, n-s resol, format, compressed -- read by
 * position, GisRasterHdr.C:26-60) and .../fcell/<name>: big-endian float32
 * rows, raw or compressed (byte 0 = 4 or 8, the size of the rows+1 big-endian
 * row offsets that follow; a row starting with '1' is zlib-deflated, any
 * other flag byte stores it raw; GisBinFile.C:131-228).  Fills dem (nrows,
 * ncols, res = e-w resolution, xmin = west, ymax = south + rows * n-s
 * resolution as set_from_header does, GisApi.C:1455-1480; dem->elev = elev)
 * and, when capacity >= rows * cols, elev [rows * cols] (row 0 north).
 * Returns 0; -1 bad argument; -2 capacity too small (dem holds the size);
 * -4 missing or malformed header (also: e-w and n-s resolutions differ by
 * a tenth or more); -5 missing, truncated or undecodable data. */
int tg_host_read_grass(const char* gisdbase, const char* location, const char* mapset,
                       const char* name, long long capacity, tg_dem* dem, float* elev);

/* ---- run inputs -------------------------------------------------------------
 * The reference's input files (Read_data, datread.C:36-451, written by its
 * GUI): <dir>/simulation.data, <dir>/scale.data (absent: no scaling) and
 * <dir>/frict.data, read token by token in the reference's order and
 * non-dimensionalised as Read_data does (lengths by the length scale, heights
 * by the height scale, speeds by sqrt(L g), times by sqrt(L / g)).  Flux
 * sources and discharge planes are counted and skipped (this path has no
 * mass sources).  piles [pile_cap]: the piles (tg_host_piles takes their
 * centres, radii, rotation and height).
 * Returns 0; -1 bad argument; -2 more piles than pile_cap (out->npiles holds
 * the count); -4 simulation.data or frict.data missing; -5 malformed. */
#define TG_HOST_MAX_MATERIALS 16
typedef struct tg_pile {
  double height, xc, yc, major, minor;  /* non-dimensional                     */
  double rot_deg, cosrot, sinrot;       /* counter-clockwise rotation          */
  double vx, vy;                        /* initial velocity, non-dimensional   */
} tg_pile;
typedef struct tg_run_inputs {
  int srctype, npiles, nsources;
  double length_scale, height_scale, gravity_scale;
  double epsilon;          /* height / length scale                           */
  double time_scale;       /* sqrt(length / gravity scale)                    */
  double smallest_axis;    /* 2 x the smallest pile radius (non-dimensional)  */
  int cells_across;        /* cells across the smallest pile axis             */
  int maxiter;
  double maxtime, timeoutput, timesave;          /* as read (seconds)          */
  double nd_maxtime, nd_nextoutput, nd_nextsave;  /* non-dimensional (tg_params) */
  int adaptflag, vizflag, order;
  char gis_main[300], gis_sub[100], gis_mapset[100], gis_map[100];
  int extramaps;
  double edge_height, test_height, test_x, test_y;  /* non-dimensional; no test
                                                      point: HUGE_VAL          */
  int nplanes;
  int nmat;                /* materials in frict.data                         */
  char matnames[TG_HOST_MAX_MATERIALS][64];
  double intfrict;         /* internal friction angle (radians), material 1's */
  double bedfrict[TG_HOST_MAX_MATERIALS];  /* bed friction angles (radians) of
                                              materials 1..nmat at [0..nmat-1] */
} tg_run_inputs;
int tg_host_read_run_inputs(const char* dir, int pile_cap, tg_pile* piles, tg_run_inputs* out);

/* ---- the grid file -----------------------------------------------------------
 * funky%04d.inp (replaces Read_grid, datread.C:460-660; the format its
 * preprocessor writes, preprocess.C:560-660, node.C:82-93,
 * element2.C:233-283): binary in the writer's byte order; version 20061109
 * (ranges and coordinates as float) or 20061110 (as double).  Node
 * coordinates and the x / y ranges are divided by length_scale, as Read_grid
 * does.  Per element: the keys of its 9 nodes (0-3 corners SW SE NE NW, 4-7
 * side nodes, 8 the bubble node = the element's key), per side 0..3 the
 * neighbour's process (-1: domain boundary) and key, the boundary
 * conditions (Read_grid reads and discards them: its BC object is not
 * attached to the element), elm_loc (the element's i, j in the base grid),
 * the opposite brother's key and the material index.  The material names
 * that follow are not read (Read_grid does not either).
 * node_keys [2*node_cap], node_xy [2*node_cap] (key / x, y interleaved).
 * Returns 0; -1 bad argument; -2 a capacity too small (hdr holds the counts);
 * -4 file missing; -5 unknown version, truncated or malformed. */
typedef struct tg_funky_header {
  int version, nnodes, nelems;
  double keyrange[2], xrange[2], yrange[2];
} tg_funky_header;
typedef struct tg_funky_elem {
  uint32_t node_key[9][2];
  int32_t neigh_proc[4];
  uint32_t neigh_key[4][2];
  int32_t bc_type[4];       /* 0 none, 1 essential, 2 natural, 3 both          */
  float bc_value[4][2][2];  /* [side][essential, natural][x, y]                */
  int32_t elm_loc[2];
  uint32_t opposite_brother[2];
  int32_t material;
} tg_funky_elem;
int tg_host_read_funky(const char* path, double length_scale, tg_funky_header* hdr,
                       int node_cap, uint32_t* node_keys, double* node_xy, int elem_cap,
                       tg_funky_elem* elems);

/* Analytic DEM: a plane of slope angle slope_deg falling towards +x that
 * ends at x_toe and blends into a flat runout with a C1 cosine transition of
 * total width blend (no reference counterpart; synthetic input). */
int tg_host_dem_plane_runout(int nrows, int ncols, double xmin, double ymax, double res,
                             double slope_deg, double x_toe, double blend, float* elev);

/* Seeded diamond-square fractal DEM on a (2^k+1)^2 grid, rescaled so that
 * max - min == relief; roughness H (amplitude falls by 2^-H per level). */
int tg_host_dem_diamond_square(int n, double relief, double H, uint64_t seed, float* elev);

/* Per-cell static coefficients of a uniform block of cells from the DEM:
 * Horn 3x3 slope grids (calculate_slope, GisApi.C:1295-1345), bilinear slope
 * sampling (Get_slope, GisApi.C:731-762), curvature (Get_curvature,
 * GisApi.C:764-833), gravity vector (calc_gravity_vector, element2.C:3545-3564)
 * and d gravity_z (calc_d_gravity, element2.C:3574-3653).
 * grav [3*n], dgrav [2*n], curv [2*n] FP64 SoA; n = nx*ny. */
int tg_host_static_coefs(const tg_dem* dem, int nx, int ny, double dx, double dy,
                         double x0, double y0, double length_scale, double gravity_scale,
                         double* grav, double* dgrav, double* curv);

/* Gravity vector and curvature of the leaves of an adaptive mesh (or any
 * list of points) at their centres xc [2n] (x, y interleaved): the
 * per-element set-up of tg_host_static_coefs (calc_topo_data,
 * element2.C:3655-3703; calc_gravity_vector :3545-3564) at arbitrary centres,
 * as the reference evaluates it for every new son and father
 * (element2.C:238-240, 524-526).  grav [3n], curv [2n] FP64 SoA; d gravity_z
 * comes from tg_host_leaf_d_gravity on the flattened table.
 * Returns 0; -1 bad argument; -2 a centre outside the raster. */
int tg_host_leaf_static_coefs(const tg_dem* dem, int n, const double* xc, double length_scale,
                              double gravity_scale, double* grav, double* curv);

/* Elliptical paraboloid piles (elliptical_pile_height, init_piles.C:167-277,
 * PARABALOID branch): each cell takes the 9-node weighted average
 * (corners + 2*edges + 4*centre)/16 of max-over-piles heights. Rotation in
 * degrees (cos/sin of rot evaluated on the host).  hu, hv are set to zero
 * (piles start at rest). */
int tg_host_piles(int nx, int ny, double dx, double dy, double x0, double y0,
                  int npiles, const double* xcen, const double* ycen,
                  const double* major, const double* minor, const double* rot_deg,
                  const double* height, double* h, double* hu, double* hv);

/* Terminal functional weights for "h integrated over a box": w_h = dx*dy for
 * cells whose centre lies strictly inside [xmin,xmax]x[ymin,ymax]
 * (box test as MaxH_Functional::clac_func_sens, calc_adjoint.C:467-491),
 * w_hu = w_hv = 0.  w is FP64 SoA [3*n]. */
int tg_host_box_weights(int nx, int ny, double dx, double dy, double x0, double y0,
                        double xmin, double xmax, double ymin, double ymax, double* w);

/* ---- adaptive-mesh flattening ------------------------------------------- */

/* Hilbert keys of normalised coordinates: replaces fhsfc2d_ (hilbert.C:267-278
 * with hsfc2d :76-145, two 32-bit words, most significant word first).
 * norm_xy [2*n] interleaved (x, y) in [0, 1]; keys [2*n] interleaved. */
int tg_host_hsfc2d(int n, const double* norm_xy, uint32_t* keys);

/* Element-table bucket of each key: HashTable::hash (hashtab.h:181-189) with
 * hashconstant = 8*nbuckets/(keyrange0*keyrange1 + keyrange1)
 * (hashtab2.C:85). */
int tg_host_hash_bucket(int n, const uint32_t* keys, double keyrange0, double keyrange1,
                        int nbuckets, int32_t* bucket);

/* Element-table traversal order of n keys added in index order: buckets
 * 0..nbuckets-1 (tg_host_hash_bucket), each bucket's list as
 * HashTable::addElement (hashtab2.C:185-217) builds it -- entries walked past
 * while the new first key word is larger; on an equal first word, walked past
 * while the second word is larger, without re-checking the first word.
 * order[p] = index of the p-th entry visited. */
int tg_host_traversal(int n, const uint32_t* keys, double keyrange0, double keyrange1,
                      int nbuckets, int32_t* order);

/* Flattened leaf tables, in the reference's element-table traversal order
 * (buckets 0..nbuckets-1, each bucket's list in HashTable::addElement order,
 * hashtab2.C:185-217).  Caller-owned arrays of the sizes given. */
typedef struct tg_leaf_table {
  int n;             /* number of leaves (set by tg_host_flatten)              */
  int32_t* order;    /* [n]   flattened position -> input leaf index            */
  uint32_t* key;     /* [2n]  element key (Element::key)                       */
  int32_t* gen;      /* [n]   refinement generation (Element::generation)      */
  int32_t* nbr;      /* [8n]  flattened index of the neighbour in slot s; -1 on
                        the domain boundary (Element::neighbor[8])             */
  int32_t* nkind;    /* [8n]  0 neighbour present, -1 domain boundary, -2 no
                        second neighbour on that side (Element::neigh_proc
                        with one process)                                      */
  int32_t* ngen;     /* [8n]  neighbour generation (Element::neigh_gen)        */
  double* dx;        /* [2n]  cell size                                        */
  double* xc;        /* [2n]  cell centre                                      */
} tg_leaf_table;

/* Flattens a 2:1-balanced quadtree over an nx0 x ny0 base grid of cells
 * (dx0, dy0) with lower-left corner (xmin, ymin).  Leaf e is the cell
 * (ix[e], iy[e]) of level lev[e] (cell size dx0/2^lev); e is also the order
 * in which the leaves were added to the element table (it decides the order
 * inside a bucket).  Side s: 0 south, 1 east, 2 north, 3 west; with two finer
 * neighbours on side s, slot s holds the one next to corner s and slot s+4
 * the one next to corner s+1 (corners 0 SW, 1 SE, 2 NE, 3 NW), as the son
 * construction in refine2.C:589-700 assigns them.  Element keys are the
 * Hilbert keys of the centres normalised by the domain extent, with key
 * ranges 2^32+1 (preprocess.C:561-562, one process).
 * Returns 0; -1 bad argument; -2 overlapping / out-of-range leaves; -3 leaves
 * that do not cover the base grid; -4 neighbours more than one level apart. */
int tg_host_flatten(int nx0, int ny0, double xmin, double ymin, double dx0, double dy0, int n,
                    const int32_t* lev, const int32_t* ix, const int32_t* iy, int nbuckets,
                    tg_leaf_table* out);

/* d gravity_z of the leaves of a flattened table (calc_d_gravity,
 * element2.C:3574-3653): per direction the weighted average of the one-sided
 * differences of gz towards the neighbours of slots 1 (x+) and 3 (x-), 2 (y+)
 * and 0 (y-) -- on a side with two finer leaves the one in slot s -- over the
 * distances between centres; one-sided where a side has no neighbour (-1),
 * 0 with none.  nbr [8n] and xc [2n] as tg_leaf_table; gz [n] the leaves'
 * gravity_z; dgrav FP64 SoA [2n].  Returns 0; -1 bad argument. */
int tg_host_leaf_d_gravity(int n, const int32_t* nbr, const double* xc, const double* gz,
                           double* dgrav);

/* ---- mesh adaptation ------------------------------------------------------ */

/* One adaptation pass (H_adapt then unrefine, as hpfem.C:244-258 runs them
 * every few steps) of a 2:1-balanced quadtree over an nx0 x ny0 base grid,
 * given by its leaves (leaf e = cell (ix[e], iy[e]) of level lev[e]) and
 * their state:
 *  - refinement (hadpt.C:50-323): every leaf below max_level on one of the
 *    pile contours h = tiny, 5 tiny, 15 tiny, 40 tiny (if_pile_boundary,
 *    element2.C:4020-4065) -- or, with el_error, whose error exceeds the mean
 *    error of the leaves with hu > tiny (element_weight.C:36-79) -- is split,
 *    coarser leaves across its sides first (2:1 cascade); then
 *    num_buffer_layer buffer layers around the contours (if_first_ and
 *    if_next_buffer_boundary, element2.C:4136-4220), refining where a
 *    buffer leaf is not yet at max_level;
 *  - unrefinement (unrefine.C:54-352): four brothers that were neither split
 *    nor made buffer in this pass and none of which has a finer neighbour
 *    become their father.
 * A son takes its father's state (element2.C:246-250), a father the sum of
 * 0.25 x its sons' states in son order (element2.C:527-532).  Sons of a leaf
 * are numbered from its corners: 0 SW, 1 SE, 2 NE, 3 NW.  No mass sources.
 * Output: the new leaves (the input order, a split leaf replaced by its sons,
 * a father at its son 0's place) and state, and the operations in the order
 * applied (for tg_host_adapt_forward / tg_host_adapt_adjoint).
 * Returns 0; -1 bad argument; -2 cap / op_cap too small (*n_out, *nops give
 * the sizes needed); -3 the input leaves do not form a 2:1-balanced tree. */
typedef struct tg_adapt_params {
  int max_level;         /* REFINE_LEVEL (datread.C:693-696), at most 28    */
  int num_buffer_layer;  /* 5 in the reference's time loop (hpfem.C:247)   */
  double tiny;           /* GEOFLOW_TINY                                   */
} tg_adapt_params;
/* kind 0: leaf (lev, ix, iy) split into its four sons; kind 1: the four
 * sons of (lev, ix, iy) merged into it. */
typedef struct tg_adapt_op {
  int32_t kind, lev, ix, iy;
} tg_adapt_op;
int tg_host_adapt(int nx0, int ny0, const tg_adapt_params* prm, int n, const int32_t* lev,
                  const int32_t* ix, const int32_t* iy, const double* h, const double* hu,
                  const double* hv, const double* el_error, int cap, int* n_out,
                  int32_t* lev_out, int32_t* ix_out, int32_t* iy_out, double* h_out,
                  double* hu_out, double* hv_out, int op_cap, int* nops, tg_adapt_op* ops);

/* The state transfer of a pass applied to any fields: f_old [nf * n_old]
 * (nf SoA fields on the old leaves) -> f_new [nf * n_new] on the new leaves,
 * with the rules of tg_host_adapt.  Returns 0; -1 bad argument; -3 the
 * leaves and the operations do not match. */
int tg_host_adapt_forward(int nops, const tg_adapt_op* ops, int nf, int n_old,
                          const int32_t* lev_old, const int32_t* ix_old, const int32_t* iy_old,
                          const double* f_old, int n_new, const int32_t* lev_new,
                          const int32_t* ix_new, const int32_t* iy_new, double* f_new);

/* Its transpose, the adjoint transfer of the reverse sweep across the pass
 * (dual_refine_unrefine.C:20-450 with the DualElem constructors,
 * dualmesh.C:708-709, 986-993): from lambda on the new leaves (SoA [3 n_new])
 * to lambda on the old ones -- a split leaf takes the sum of its sons'
 * adjoints in son order, each son of a merge 0.25 of its father's -- with
 * the operations undone in reverse order.  Returns 0; -1 bad argument; -3
 * mismatch. */
int tg_host_adapt_adjoint(int nops, const tg_adapt_op* ops, int n_new, const int32_t* lev_new,
                          const int32_t* ix_new, const int32_t* iy_new, const double* lam_new,
                          int n_old, const int32_t* lev_old, const int32_t* ix_old,
                          const int32_t* iy_old, double* lam_old);

/* Trajectory spill file of one step (SolRec::wrtie_sol_to_disk,
 * dualmesh.C:87-142; the reference names it solution_<rank %04d>_<step %08d>):
 * a gzip stream of int32 nzero, int32 nnz, nzero element keys (2 x uint32),
 * then nnz records of key (2 x uint32) and h, hu, hv (FP64), host byte
 * order.  Elements are taken in `order` (positions into the arrays; NULL =
 * 0..n-1); an element whose h > 0 is false (h <= 0 or NaN) is a zero entry
 * (SolRec::record_solution, dualmesh.C:53-70).
 * Returns 0; -1 bad argument; -2 I/O error. */
int tg_host_write_solution(const char* path, int64_t n, const uint32_t* keys,
                           const int32_t* order, const double* h, const double* hu,
                           const double* hv);

/* Reads such a file into the elements with the given keys
 * (SolRec::read_sol_from_disk, dualmesh.C:234-297): a zero entry gives
 * (0, 0, 0), an element absent from the file keeps its values.  found: the
 * file's entries matched to a key; unknown: those whose key is not among
 * `keys` (the reference adds them to its table; here they are counted).
 * Returns 0; -1 bad argument; -2 cannot open; -3 truncated or trailing data. */
int tg_host_read_solution(const char* path, int64_t n, const uint32_t* keys, double* h,
                          double* hu, double* hv, int64_t* found, int64_t* unknown);

/* ---- error grid ------------------------------------------------------------
 * The reference's error grid is the mesh with every leaf refined once
 * (uinform_refine, uniform_refine.C:24-130); its elements carry their
 * father's state, previous state and adjoint, and bilinear_interp
 * (bilinear_interp.C:126-556) reconstructs each from a stencil of four
 * neighbouring elements.  tg_host_error_stencil picks every error element's
 * stencil as bilinear_interp does and returns the rule and the points:
 *   kind[e]:  TG_ERR_INTERP_COPY (one distinct point: the element keeps its
 *             father's values), _LINEAR_X / _LINEAR_Y (two points, sorted by
 *             x then y: linear_interp along that axis), _BARYCENTRIC (three),
 *             _BILINEAR (four: the 4x4 fit of a0 + a1 x + a2 y + a3 x y);
 *   pts[4e + k]: father leaf of point k (-1 unused); pxy[8e + 2k], [8e + 2k + 1]:
 *             its x, y (the father's centre as the reference computes it:
 *             the element's centre +- half its size by which_son).
 * Inputs: the error grid's tables as tg_host_flatten returns them (nbr,
 * nproc [8n], dx, xc [2n]), which_son[e] (0 SW, 1 SE, 2 NE, 3 NW) and the
 * father leaf father[e] in 0 .. nfather-1 (every father has four sons).
 * Returns 0; -1 bad argument; -2 sons and fathers do not match. */
enum {
  TG_ERR_INTERP_COPY = 0,
  TG_ERR_INTERP_LINEAR_X = 1,
  TG_ERR_INTERP_LINEAR_Y = 2,
  TG_ERR_INTERP_BARYCENTRIC = 3,
  TG_ERR_INTERP_BILINEAR = 4
};
int tg_host_error_stencil(int n, const int32_t* nbr, const int32_t* nproc, const double* cdx,
                          const double* xc, const int32_t* which_son, const int32_t* father,
                          int nfather, int32_t* kind, int32_t* pts, double* pxy);

#ifdef __cplusplus
}
#endif
#endif
