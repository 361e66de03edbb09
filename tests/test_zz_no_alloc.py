"""GPU: tg_step_noalloc (one forward step with every intermediate in the
library's fixed device array, nothing allocated) against tg_forward_step on
the same states.  Bar: bit-identical states and dt at every step, over a run
long enough to cross the frict_tiny switch, then a second, smaller block in
the same process (the fixed array keeps the last call's contents)."""
import ctypes as C
import sys

import numpy as np
import pytest

from discrete_adjoint_amd import _lib
from discrete_adjoint_amd.solver import Simulation
from helpers import small_flow_problem
from hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _reference(pb, nsteps):
    """States and dt of tg_forward_step, step by step."""
    out = []
    with Simulation(pb, max_steps=nsteps) as sim:
        cfg, prm = sim.cfg, sim.params
        for _ in range(nsteps):
            st = sim.step()
            out.append((sim.get_state(), st.dt, st.time))
    return cfg, prm, out


def _noalloc_chain(pb, cfg, prm, nsteps):
    lib = _lib.load()
    statics = DeviceArray.of(np.concatenate([pb.grav.reshape(-1), pb.dgrav.reshape(-1),
                                             pb.curv.reshape(-1)]).astype(np.float64))
    bufs = [[DeviceArray.of(np.asarray(a, dtype=np.float64)) for a in (pb.h, pb.hu, pb.hv)],
            [DeviceArray(pb.n) for _ in range(3)]]
    blk = _lib.tg_noalloc_block(pb.nx, pb.ny, pb.dx, pb.dy, statics.ptr)
    synchronize()
    out, t = [], 0.0
    try:
        for s in range(nsteps):
            src, dst = bufs[s % 2], bufs[1 - s % 2]
            dt = C.c_double()
            rc = lib.tg_step_noalloc(C.byref(cfg), C.byref(blk), C.byref(prm), s, t,
                                     *[b.ptr for b in src], *[b.ptr for b in dst], None,
                                     C.byref(dt))
            assert rc == _lib.TG_OK, rc
            t += dt.value
            out.append(([b.numpy() for b in dst], dt.value))
        synchronize()
    finally:
        for b in [statics] + bufs[0] + bufs[1]:
            b.free()
    return out


@pytest.mark.parametrize("nx,ny,nsteps", [(300, 210, 60), (70, 45, 12)])
def test_noalloc_step_equals_forward_step(nx, ny, nsteps):
    pb = small_flow_problem(nx, ny, nsteps)
    _say(f"case noalloc {nx}x{ny} {nsteps} steps")
    cfg, prm, ref = _reference(pb, nsteps)
    got = _noalloc_chain(pb, cfg, prm, nsteps)
    for s, ((U, dt), (Ur, dtr, _)) in enumerate(zip(got, ref)):
        assert dt == dtr, s
        for a, b in zip(U, Ur):
            assert np.array_equal(a, b), s
    assert ref[-1][0][0].max() > 0


def test_noalloc_rejects_blocks_beyond_its_scratch():
    lib = _lib.load()
    cfg = _lib.tg_config()
    lib.tg_config_default(C.byref(cfg))
    prm = _lib.tg_params()
    assert lib.tg_params_from_angles(0.5, 0.3, 1.0, 1.0, C.byref(prm)) == _lib.TG_OK
    blk = _lib.tg_noalloc_block(1025, 1024, 1.0, 1.0, 8)   # one row too many
    dt = C.c_double()
    rc = lib.tg_step_noalloc(C.byref(cfg), C.byref(blk), C.byref(prm), 0, 0.0, 8, 8, 8, 8, 8,
                             8, None, C.byref(dt))
    assert rc == _lib.TG_ERR_TOO_LARGE
    assert lib.tg_noalloc_scratch_bytes() < 16 << 20
