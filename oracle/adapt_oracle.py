"""TEST INFRASTRUCTURE ONLY: CPU restatement of the adaptive time loop and its
reverse sweep, for checking the product's (discrete_adjoint_amd.adaptive) GPU
run.  Imported by tests/ only, as the checker.

- The state transfer of a refinement pass, restated from the reference: a son
  copies its father (Element son constructor, element2.C:246-250, times
  myfractionoffather = 1), a new father sums 0.25 x its sons' states in son
  order (element2.C:527-532).
- Its adjoint, restated from the dual elements: a new dual son takes 0.25 of
  its father's adjoint (dualmesh.C:708-709), a new dual father the sum of its
  sons' adjoints in son order (dualmesh.C:986-993), undoing the passes last
  to first (dual_refine_unrefine.C:20-450 replays the forward passes
  backwards).
- The time loop: the leaf oracle's steps on each mesh between passes, a pass
  before step s when s % R == R - 1 (properties.h:482-486), the reverse
  sweep over every segment's stored states.

Which leaves split or merge is the product's host decision (tg_host_adapt,
pinned by tests/test_adapt.py); this restatement applies the operations it
returns.
"""
import numpy as np

import oracle

SONS = ((0, 0), (1, 0), (1, 1), (0, 1))   # SW, SE, NE, NW


def _cells(tree):
    lev, ix, iy = tree
    return list(zip(np.asarray(lev).tolist(), np.asarray(ix).tolist(), np.asarray(iy).tolist()))


def transfer_forward(ops, old_tree, U_old, new_tree):
    """State (3, n_old) on old_tree -> (3, n_new) on new_tree."""
    val = {c: [float(U_old[v][e]) for v in range(3)] for e, c in enumerate(_cells(old_tree))}
    for o in ops:
        L, i, j = int(o["lev"]), int(o["ix"]), int(o["iy"])
        if o["kind"] == 0:
            f = val.pop((L, i, j))
            for a, b in SONS:
                val[(L + 1, 2 * i + a, 2 * j + b)] = [x * 1.0 for x in f]
        else:
            f = [0.0, 0.0, 0.0]
            for a, b in SONS:
                s = val.pop((L + 1, 2 * i + a, 2 * j + b))
                for v in range(3):
                    f[v] += s[v] * 0.25
            val[(L, i, j)] = f
    cells = _cells(new_tree)
    assert len(val) == len(cells)
    return np.array([[val[c][v] for c in cells] for v in range(3)])


def transfer_adjoint(ops, new_tree, lam_new, old_tree):
    """Adjoint (3, n_new) on new_tree -> (3, n_old) on old_tree."""
    val = {c: [float(lam_new[v][e]) for v in range(3)] for e, c in enumerate(_cells(new_tree))}
    for o in reversed(list(ops)):
        L, i, j = int(o["lev"]), int(o["ix"]), int(o["iy"])
        if o["kind"] == 0:
            f = [0.0, 0.0, 0.0]
            for a, b in SONS:
                s = val.pop((L + 1, 2 * i + a, 2 * j + b))
                for v in range(3):
                    f[v] += s[v]
            val[(L, i, j)] = f
        else:
            f = val.pop((L, i, j))
            for a, b in SONS:
                val[(L + 1, 2 * i + a, 2 * j + b)] = [0.25 * x for x in f]
    cells = _cells(old_tree)
    assert len(val) == len(cells)
    return np.array([[val[c][v] for c in cells] for v in range(3)])


def run(driver, nsteps, weights_fn, ops_of=None, U0=None, want_dphi=True, dt_seq=None,
        phi_bed_deg=None):
    """The adaptive loop of `driver` (an adaptive.AdaptiveRun: geometry,
    statics, angles, adapt_every) on the leaf oracle.  ops_of(k, tree, U) ->
    (new tree, ops) gives pass k's operations (None: the product's decision
    on the oracle's state); dt_seq: prescribed dt of every step (None: CFL);
    U0 / phi_bed_deg: another initial state / bed friction angle.
    weights_fn(final tree) -> weights (3, n).  Returns a dict: final tree
    and state, lambda^0 on the initial tree, dJ/dphi_bed, the trees and
    operations of every pass, the dt record."""
    from discrete_adjoint_amd import mesh
    tree = driver.tree
    U = np.array(driver.U if U0 is None else U0, dtype=np.float64, copy=True)
    time = 0.0
    segs = []
    for k, (a, b) in enumerate(driver._bounds(nsteps)):
        ops = None
        prev = tree
        if a > 0:
            if ops_of is None:
                lev, ix, iy, _, p = mesh.adapt(*driver.geometry[:2], *tree, U, driver.max_level,
                                               driver.num_buffer_layer, driver.tiny)
                tree, ops = (lev, ix, iy), p.ops
            else:
                tree, ops = ops_of(k, tree, U)
            U = transfer_forward(ops, prev, U, tree)
        pb = driver.problem_of(tree, U)
        o = pb.table.order
        orc = oracle.LeafOracle.from_problem(pb)
        if phi_bed_deg is not None:
            orc.set_angles(driver.phi_int_deg, phi_bed_deg)
        orc.time = time
        Uf = [np.ascontiguousarray(U[v][o]) for v in range(3)]
        traj, dts = [], []
        for s in range(a, b):
            traj.append([x.copy() for x in Uf])
            dt, _ = orc.step(Uf, s, None if dt_seq is None else dt_seq[s])
            dts.append(dt)
        time = orc.time
        U = np.empty((3, len(o)))
        U[:, o] = np.array(Uf)
        segs.append(dict(tree=tree, prev=prev, ops=ops, orc=orc, traj=traj, dts=dts, a=a,
                         order=o))
    lam = np.array(weights_fn(tree), dtype=np.float64).reshape(3, -1)
    dphi = 0.0
    for sg in reversed(segs):
        o = sg["order"]
        lf = np.ascontiguousarray(lam[:, o])
        for i in range(len(sg["traj"]) - 1, -1, -1):
            lf, acc = sg["orc"].adjoint_step(sg["traj"][i], sg["a"] + i, sg["dts"][i], lf,
                                             want_dphi=want_dphi)
            dphi += acc
        lam = np.empty((3, len(o)))
        lam[:, o] = lf
        if sg["ops"] is not None:
            lam = transfer_adjoint(sg["ops"], sg["tree"], lam, sg["prev"])
    return dict(tree=tree, U=U, lam0=lam, dphi=dphi,
                passes=[(sg["tree"], sg["ops"]) for sg in segs],
                dts=np.concatenate([sg["dts"] for sg in segs]))

