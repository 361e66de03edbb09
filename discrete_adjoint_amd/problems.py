"""Synthetic problem set-up (the host side of the reference's driver: the
GIS raster, setup_geoflow's per-cell coefficients and init_piles), built
through the C host helpers in libtg.so.

A Problem carries everything tg_set_mesh / tg_set_state / tg_params need:
uniform block geometry, FP64 static coefficients (SoA), the initial state and
the material angles.  Units are unscaled (LENGTH = HEIGHT = GRAVITY scale 1,
epsilon = 1), as in the benchmark configurations.
"""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Problem:
    name: str
    nx: int
    ny: int
    dx: float
    dy: float
    x0: float
    y0: float
    grav: np.ndarray          # (3, n)
    dgrav: np.ndarray         # (2, n)
    curv: np.ndarray          # (2, n)
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    phi_int_deg: float
    phi_bed_deg: float
    nsteps: int
    checkpoint_every: int
    box: tuple = None         # functional box (xmin, xmax, ymin, ymax)
    epsilon: float = 1.0
    gravity_scale: float = 1.0
    meta: dict = field(default_factory=dict)
    # per-cell material index (uint8, n) and the bed friction angle (deg) of
    # each material; None: one material, phi_bed_deg
    mat: np.ndarray = None
    materials: tuple = None

    @property
    def n(self):
        return self.nx * self.ny

    def box_weights(self, box=None):
        box = box or self.box
        lib = _lib.load()
        w = np.zeros(3 * self.n)
        rc = lib.tg_host_box_weights(self.nx, self.ny, self.dx, self.dy, self.x0, self.y0,
                                     box[0], box[1], box[2], box[3], _lib.dptr(w))
        assert rc == 0
        return w


def _dem(nrows, ncols, xmin, ymax, res, elev):
    d = _lib.tg_dem()
    d.nrows, d.ncols, d.xmin, d.ymax, d.res = nrows, ncols, xmin, ymax, res
    d.elev = _lib.fptr(elev)
    return d


def read_run_inputs(directory):
    """The reference's run inputs (tg_host_read_run_inputs: simulation.data,
    scale.data, frict.data in `directory`), non-dimensionalised: returns
    (fields dict, list of pile dicts)."""
    lib = _lib.load()
    out = _lib.tg_run_inputs()
    cap = 64
    while True:
        piles = (_lib.tg_pile * cap)()
        rc = lib.tg_host_read_run_inputs(str(directory).encode(), cap, piles, C.byref(out))
        if rc == -2:
            cap = out.npiles
            continue
        if rc:
            raise ValueError(f"tg_host_read_run_inputs returned {rc}")
        break
    f = {}
    for name, _ in _lib.tg_run_inputs._fields_:
        v = getattr(out, name)
        if name == "matnames":
            v = [bytes(v[i].value).decode() for i in range(out.nmat)]
        elif name == "bedfrict":
            v = list(v[: out.nmat])
        elif isinstance(v, bytes):
            v = v.decode()
        f[name] = v
    pl = [{n: getattr(piles[i], n) for n, _ in _lib.tg_pile._fields_} for i in range(out.npiles)]
    return f, pl


def read_funky(path, length_scale=1.0):
    """The reference's grid file (tg_host_read_funky): returns (header dict,
    node keys (n, 2) uint32, node coordinates (n, 2), elements: dict of
    arrays node_key (m, 9, 2), neigh_proc (m, 4), neigh_key (m, 4, 2),
    bc_type (m, 4), bc_value (m, 4, 2, 2), elm_loc (m, 2), opposite_brother
    (m, 2), material (m,)); lengths divided by length_scale."""
    lib = _lib.load()
    hdr = _lib.tg_funky_header()
    rc = lib.tg_host_read_funky(str(path).encode(), float(length_scale), C.byref(hdr), 0, None,
                                None, 0, None)
    if rc not in (0, -2):
        raise ValueError(f"tg_host_read_funky returned {rc}")
    nn, ne = hdr.nnodes, hdr.nelems
    keys = np.zeros((max(nn, 1), 2), dtype=np.uint32)
    xy = np.zeros((max(nn, 1), 2))
    elems = (_lib.tg_funky_elem * max(ne, 1))()
    rc = lib.tg_host_read_funky(str(path).encode(), float(length_scale), C.byref(hdr), nn,
                                keys.ctypes.data_as(C.POINTER(C.c_uint32)), _lib.dptr(xy), ne,
                                elems)
    if rc:
        raise ValueError(f"tg_host_read_funky returned {rc}")
    raw = np.ctypeslib.as_array(elems)[:ne]
    el = {name: np.array(raw[name]) for name, _ in _lib.tg_funky_elem._fields_}
    h = {"version": hdr.version, "nnodes": nn, "nelems": ne,
         "keyrange": list(hdr.keyrange), "xrange": list(hdr.xrange), "yrange": list(hdr.yrange)}
    return h, keys[:nn], xy[:nn], el


def funky_base_grid(header, elems):
    """The base grid a grid file's elements tile (the preprocessor's
    uniform nx x ny grid over the x / y ranges, createfunky.C:250-345):
    (nx0, ny0, xmin, ymin, dx0, dy0, material (ny0, nx0) with -1 where no
    element of this file lies)."""
    loc = elems["elm_loc"]
    nx0, ny0 = int(loc[:, 0].max()) + 1, int(loc[:, 1].max()) + 1
    xmin, xmax = header["xrange"]
    ymin, ymax = header["yrange"]
    mat = np.full((ny0, nx0), -1, dtype=np.int32)
    mat[loc[:, 1], loc[:, 0]] = elems["material"]
    return nx0, ny0, xmin, ymin, (xmax - xmin) / nx0, (ymax - ymin) / ny0, mat


def read_grass(gisdbase, location, mapset, name):
    """A GRASS elevation raster (tg_host_read_grass) as the DEM arguments
    (nrows, ncols, xmin, ymax, res, elev float32) of static_coefs."""
    lib = _lib.load()
    d = _lib.tg_dem()
    args = [str(a).encode() for a in (gisdbase, location, mapset, name)]
    rc = lib.tg_host_read_grass(*args, 0, C.byref(d), None)
    if rc not in (0, -2):
        raise ValueError(f"tg_host_read_grass returned {rc}")
    elev = np.empty(d.nrows * d.ncols, dtype=np.float32)
    rc = lib.tg_host_read_grass(*args, len(elev), C.byref(d), _lib.fptr(elev))
    if rc:
        raise ValueError(f"tg_host_read_grass returned {rc}")
    return (d.nrows, d.ncols, d.xmin, d.ymax, d.res, elev)


def static_coefs(dem_args, nx, ny, dx, dy, x0, y0):
    lib = _lib.load()
    n = nx * ny
    grav = np.zeros(3 * n)
    dgrav = np.zeros(2 * n)
    curv = np.zeros(2 * n)
    dem = _dem(*dem_args)
    rc = lib.tg_host_static_coefs(C.byref(dem), nx, ny, dx, dy, x0, y0, 1.0, 1.0,
                                  _lib.dptr(grav), _lib.dptr(dgrav), _lib.dptr(curv))
    if rc != 0:
        raise ValueError(f"tg_host_static_coefs failed ({rc})")
    return grav.reshape(3, n), dgrav.reshape(2, n), curv.reshape(2, n)


def piles(nx, ny, dx, dy, x0, y0, specs):
    """specs: list of (xc, yc, major, minor, rot_deg, height)."""
    lib = _lib.load()
    n = nx * ny
    a = np.array(specs, dtype=np.float64).reshape(-1, 6)
    cols = [np.ascontiguousarray(a[:, k]) for k in range(6)]
    h, hu, hv = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = lib.tg_host_piles(nx, ny, dx, dy, x0, y0, len(specs), *[_lib.dptr(c) for c in cols],
                           _lib.dptr(h), _lib.dptr(hu), _lib.dptr(hv))
    assert rc == 0
    return h, hu, hv


def paraboloid_radius_for_volume(volume, aspect=0.1):
    """Circular paraboloid with height = aspect * radius: V = pi/2 r^2 H."""
    return (2.0 * volume / (math.pi * aspect)) ** (1.0 / 3.0)


def fractal_dem(n=1025, relief=800.0, H=0.8, seed=1):
    lib = _lib.load()
    elev = np.zeros(n * n, dtype=np.float32)
    rc = lib.tg_host_dem_diamond_square(n, relief, H, seed, _lib.fptr(elev))
    assert rc == 0
    return elev


def plane_runout_problem(n=128, dx=10.0, nsteps=200):
    """Uniform n x n grid of dx cells on an analytic DEM: a 35-degree plane
    over 60% of the x extent with a C1 blend (100 m) into a flat runout; one
    paraboloid pile 150 x 100 m, 30 m high, centred on the slope; phi_int 30,
    phi_bed 20 degrees; functional box 200 x 200 m on the runout."""
    L = n * dx
    res = dx / 2.0
    margin = 4 * res
    xmin, ymax = -margin, L + margin
    ncols = nrows = int(round((L + 2 * margin) / res)) + 1
    elev = np.zeros(nrows * ncols, dtype=np.float32)
    lib = _lib.load()
    rc = lib.tg_host_dem_plane_runout(nrows, ncols, xmin, ymax, res, 35.0, 0.6 * L, 100.0,
                                      _lib.fptr(elev))
    assert rc == 0
    x0 = y0 = 0.5 * dx
    grav, dgrav, curv = static_coefs((nrows, ncols, xmin, ymax, res, elev), n, n, dx, dx, x0, y0)
    h, hu, hv = piles(n, n, dx, dx, x0, y0, [(0.3 * L, 0.5 * L, 150.0, 100.0, 0.0, 30.0)])
    xb = 0.6 * L + 100.0
    box = (xb, xb + 200.0, 0.5 * L - 100.0, 0.5 * L + 100.0)
    return Problem("plane_runout_%d" % n, n, n, dx, dx, x0, y0, grav, dgrav, curv, h, hu, hv,
                   30.0, 20.0, nsteps, 50, box=box)


def fractal_problem(n=512, dx=5.0, nsteps=1000, npiles=1, seed=3, dem_seed=1,
                    volumes=None, phi_int=32.0, phi_bed=22.0, checkpoint_every=50):
    """Uniform n x n grid on a window of a seeded 1025^2 diamond-square DEM
    (5 m spacing, 800 m relief, H = 0.8); piles at seeded positions."""
    dem_n, dem_res = 1025, 5.0
    elev = fractal_dem(dem_n, 800.0, 0.8, dem_seed)
    span = (dem_n - 1) * dem_res            # 5120 m
    L = n * dx
    if L > span:
        raise ValueError("domain larger than the DEM")
    off = 0.5 * (span - L)
    x0 = y0 = off + 0.5 * dx
    grav, dgrav, curv = static_coefs((dem_n, dem_n, 0.0, span, dem_res, elev), n, n, dx, dx,
                                     x0, y0)
    rng = np.random.default_rng(seed)
    if volumes is None:
        volumes = list(np.exp(rng.uniform(math.log(1e5), math.log(1e6), npiles)))
    specs = []
    for v in volumes:
        r = paraboloid_radius_for_volume(v)
        cx = off + rng.uniform(0.3, 0.7) * L
        cy = off + rng.uniform(0.3, 0.7) * L
        specs.append((cx, cy, r, r, 0.0, 0.1 * r))
    h, hu, hv = piles(n, n, dx, dx, x0, y0, specs)
    p = Problem("fractal_%d" % n, n, n, dx, dx, x0, y0, grav, dgrav, curv, h, hu, hv,
                phi_int, phi_bed, nsteps, checkpoint_every)
    p.meta["piles"] = specs
    p.meta["seed"] = seed
    return p


def ensemble_problems(count=256, seed=2024, nsteps=500, sides=(64, 128, 256, 512), dx=5.0,
                      checkpoint_every=25):
    """A stochastic-study ensemble: each member a uniform side x side block
    (side drawn from `sides`) of dx cells on its own seeded diamond-square DEM
    (relief 0.156 * extent, H = 0.8), phi_bed ~ U(15, 30) deg,
    phi_int ~ U(28, 38) deg, one paraboloid pile of volume
    lognormal(ln 3e5 m^3, 0.7) at a seeded position.  The functional box is
    the middle half of the block (h integrated at the final step)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        side = int(rng.choice(sides))
        L = side * dx
        elev = fractal_dem(side + 1, 0.15625 * L, 0.8, int(rng.integers(1, 2**62)))
        x0 = y0 = 0.5 * dx
        grav, dgrav, curv = static_coefs((side + 1, side + 1, 0.0, L, dx, elev), side, side,
                                         dx, dx, x0, y0)
        vol = float(np.exp(rng.normal(math.log(3e5), 0.7)))
        r = paraboloid_radius_for_volume(vol)
        cx, cy = rng.uniform(0.35, 0.65) * L, rng.uniform(0.35, 0.65) * L
        h, hu, hv = piles(side, side, dx, dx, x0, y0, [(cx, cy, r, r, 0.0, 0.1 * r)])
        pb = Problem("member_%d" % b, side, side, dx, dx, x0, y0, grav, dgrav, curv, h, hu, hv,
                     float(rng.uniform(28.0, 38.0)), float(rng.uniform(15.0, 30.0)), nsteps,
                     checkpoint_every, box=(0.25 * L, 0.75 * L, 0.25 * L, 0.75 * L))
        pb.meta["volume"] = vol
        out.append(pb)
    return out


def target_box(problem, h_final, cells=20, seed=5):
    """Seeded cells x cells target region placed on the final deposit: the
    window centre is a seeded draw among cells with h above 10% of max."""
    rng = np.random.default_rng(seed)
    hh = h_final.reshape(problem.ny, problem.nx)
    cand = np.argwhere(hh > 0.1 * hh.max())
    j, i = cand[rng.integers(len(cand))]
    half = cells // 2
    i0 = int(min(max(i - half, 0), problem.nx - cells))
    j0 = int(min(max(j - half, 0), problem.ny - cells))
    xa = problem.x0 + (i0 - 0.5) * problem.dx
    ya = problem.y0 + (j0 - 0.5) * problem.dy
    return (xa, xa + cells * problem.dx, ya, ya + cells * problem.dy)


@dataclass
class LeafProblem:
    """An adaptive (2:1-balanced) mesh problem: the leaf tables of
    mesh.flatten (the element-table traversal order), per-leaf FP64 static
    coefficients, the initial state and the material angles."""
    name: str
    nx0: int
    ny0: int
    dx0: float
    dy0: float
    xmin: float
    ymin: float
    table: object             # mesh.LeafTable
    grav: np.ndarray          # (3, n)
    dgrav: np.ndarray         # (2, n)
    curv: np.ndarray          # (2, n)
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    phi_int_deg: float
    phi_bed_deg: float
    nsteps: int
    checkpoint_every: int
    box: tuple = None
    epsilon: float = 1.0
    gravity_scale: float = 1.0
    meta: dict = field(default_factory=dict)
    mat: np.ndarray = None    # as Problem
    materials: tuple = None

    @property
    def n(self):
        return len(self.table.gen)

    def box_weights(self, box=None):
        """Terminal weights of sum over leaves with centre inside the box of
        h * dx * dy (the uniform tg_host_box_weights rule, per leaf)."""
        box = box or self.box
        x, y = self.table.xc[:, 0], self.table.xc[:, 1]
        inside = (x > box[0]) & (x < box[1]) & (y > box[2]) & (y < box[3])
        w = np.zeros((3, self.n))
        w[0, inside] = (self.table.dx[:, 0] * self.table.dx[:, 1])[inside]
        return w.reshape(-1)


def adaptive_problem(nx0=32, ny0=32, dx0=40.0, max_level=3, nsteps=200, refine_radius=1.6,
                     pile=None, dem=None, phi_int=30.0, phi_bed=20.0, checkpoint_every=50,
                     nbuckets=None, name=None):
    """Adaptive mesh around one paraboloid pile: base nx0 x ny0 cells of dx0,
    cells whose centre lies within refine_radius * pile radius of the pile
    centre refined down to max_level, then 2:1-balanced and flattened into the
    element-table order.  dem: None for the 35-degree plane / runout DEM of
    plane_runout_problem over the same extent, or a (nrows, ncols, xmin, ymax,
    res, elev) raster.  Gravity and curvature of a leaf are those of its own
    level's uniform grid (tg_host_static_coefs at that level's spacing), d
    gravity_z comes from the leaf's own neighbours (tg_host_leaf_d_gravity);
    the initial state samples the paraboloid at the leaf centres."""
    from . import mesh as _mesh
    Lx, Ly = nx0 * dx0, ny0 * dx0
    if pile is None:
        pile = (0.3 * Lx, 0.5 * Ly, 150.0, 100.0, 0.0, 30.0)
    cx, cy, a, b = pile[0], pile[1], pile[2], pile[3]
    rr = refine_radius * max(a, b)

    def want(L, I, J):
        s = dx0 / (1 << L)
        X, Y = (I + 0.5) * s, (J + 0.5) * s
        return np.hypot(X - cx, Y - cy) < rr + s

    lev, ix, iy = _mesh.refine_tree(nx0, ny0, max_level, want)
    nb = nbuckets or _mesh.EL_TABLE_SIZE
    t = _mesh.flatten(nx0, ny0, 0.0, 0.0, dx0, dx0, lev, ix, iy, nbuckets=nb)
    if dem is None:
        res = dx0 / (1 << max_level) / 2.0
        margin = 4 * dx0
        xmin, ymax = -margin, max(Lx, Ly) + margin
        ncols = nrows = int(round((max(Lx, Ly) + 2 * margin) / res)) + 1
        elev = np.zeros(nrows * ncols, dtype=np.float32)
        rc = _lib.load().tg_host_dem_plane_runout(nrows, ncols, xmin, ymax, res, 35.0, 0.6 * Lx,
                                                  100.0, _lib.fptr(elev))
        assert rc == 0
        dem = (nrows, ncols, xmin, ymax, res, elev)
    n = len(lev)
    grav, dgrav, curv = np.zeros((3, n)), np.zeros((2, n)), np.zeros((2, n))
    h, hu, hv = np.zeros(n), np.zeros(n), np.zeros(n)
    o = t.order
    for L in range(max_level + 1):
        sel = np.nonzero(lev[o] == L)[0]          # flattened positions at level L
        if len(sel) == 0:
            continue
        nxl, nyl, s = nx0 << L, ny0 << L, dx0 / (1 << L)
        g, _, cv = static_coefs(dem, nxl, nyl, s, s, 0.5 * s, 0.5 * s)
        hh, hhu, hhv = piles(nxl, nyl, s, s, 0.5 * s, 0.5 * s, [pile])
        cell = iy[o[sel]].astype(np.int64) * nxl + ix[o[sel]]
        grav[:, sel], curv[:, sel] = g[:, cell], cv[:, cell]
        h[sel], hu[sel], hv[sel] = hh[cell], hhu[cell], hhv[cell]
    dgrav = _mesh.leaf_d_gravity(t, grav[2])  # across the leaves' own neighbours
    xb = 0.6 * Lx + 100.0
    box = (xb, xb + 200.0, 0.5 * Ly - 100.0, 0.5 * Ly + 100.0)
    pb = LeafProblem(name or f"adaptive_{nx0}x{ny0}_L{max_level}", nx0, ny0, dx0, dx0, 0.0, 0.0,
                     t, grav, dgrav, curv, h, hu, hv, phi_int, phi_bed, nsteps, checkpoint_every,
                     box=box)
    pb.meta["levels"] = np.bincount(lev, minlength=max_level + 1).tolist()
    return pb


def adaptive_fractal_problem(nx0=128, dx0=40.0, max_level=3, radius=2600.0, npiles=2, seed=11,
                             dem_seed=1, nsteps=200, phi_int=32.0, phi_bed=22.0,
                             checkpoint_every=25):
    """A fixed adaptive mesh on the seeded 1025^2 diamond-square DEM (5 m,
    800 m relief, H = 0.8): base nx0 x nx0 cells of dx0 over the middle of
    the raster, refined to max_level (dx0 / 2^max_level) within `radius` of
    any pile centre, 2:1-balanced and flattened into the element-table
    order; npiles paraboloid piles of log-uniform volume in [1e5, 1e6] m^3
    at seeded positions.  The defaults give 0.88 M leaves (0.878 M of them
    at 5 m around the piles, hanging sides on every level).  Static coefficients of a leaf are those of its own
    level's uniform grid; the functional box is the 400 m square centred
    on the first pile."""
    from . import mesh as _mesh
    dem_n, dem_res = 1025, 5.0
    elev = fractal_dem(dem_n, 800.0, 0.8, dem_seed)
    span = (dem_n - 1) * dem_res
    L = nx0 * dx0
    if L > span:
        raise ValueError("domain larger than the DEM")
    off = 0.5 * (span - L)
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(npiles):
        v = float(np.exp(rng.uniform(math.log(1e5), math.log(1e6))))
        r = paraboloid_radius_for_volume(v)
        specs.append((rng.uniform(0.3, 0.7) * L, rng.uniform(0.3, 0.7) * L, r, r, 0.0, 0.1 * r))

    def want(Lv, I, J):
        s = dx0 / (1 << Lv)
        X, Y = (I + 0.5) * s, (J + 0.5) * s
        near = np.zeros(X.shape, bool)
        for (cx, cy, *_r) in specs:
            near |= np.hypot(X - cx, Y - cy) < radius + s
        return near

    lev, ix, iy = _mesh.refine_tree(nx0, nx0, max_level, want)
    t = _mesh.flatten(nx0, nx0, 0.0, 0.0, dx0, dx0, lev, ix, iy)
    n = len(lev)
    grav, dgrav, curv = np.zeros((3, n)), np.zeros((2, n)), np.zeros((2, n))
    h, hu, hv = np.zeros(n), np.zeros(n), np.zeros(n)
    o = t.order
    dem = (dem_n, dem_n, -off, span - off, dem_res, elev)  # raster in domain coordinates
    for Lv in range(max_level + 1):
        sel = np.nonzero(lev[o] == Lv)[0]
        if len(sel) == 0:
            continue
        nl, s = nx0 << Lv, dx0 / (1 << Lv)
        g, _, cv = static_coefs(dem, nl, nl, s, s, 0.5 * s, 0.5 * s)
        hh, hhu, hhv = piles(nl, nl, s, s, 0.5 * s, 0.5 * s, specs)
        cell = iy[o[sel]].astype(np.int64) * nl + ix[o[sel]]
        grav[:, sel], curv[:, sel] = g[:, cell], cv[:, cell]
        h[sel], hu[sel], hv[sel] = hh[cell], hhu[cell], hhv[cell]
    dgrav = _mesh.leaf_d_gravity(t, grav[2])
    cx, cy = specs[0][0], specs[0][1]
    box = (cx - 200.0, cx + 200.0, cy - 200.0, cy + 200.0)
    pb = LeafProblem(f"adaptive_fractal_{n}", nx0, nx0, dx0, dx0, 0.0, 0.0, t, grav, dgrav, curv,
                     h, hu, hv, phi_int, phi_bed, nsteps, checkpoint_every, box=box)
    pb.meta["levels"] = np.bincount(lev, minlength=max_level + 1).tolist()
    pb.meta["piles"] = specs
    pb.meta["dem"] = dem
    return pb


def leaf_dem_statics(pb):
    """statics(xc) -> (grav, curv) at arbitrary centres on the raster a leaf
    problem was built on (tg_host_leaf_static_coefs): the set-up of new
    sons, e.g. of an error grid."""
    from .adaptive import dem_statics
    return dem_statics(pb.meta["dem"])
