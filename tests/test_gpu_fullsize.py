"""GPU: the benchmark workloads at their full sizes, through properties that do not depend on size.

At these sizes the FP64 oracle would take hours, so the checks are the ones the smaller parity tests pin
against it, carried to full size: the default multi-step kernels must give the bits of one step per launch
(fields and FP64 DFT sums), and the fields must stay finite.
- 3D SOI MMI, 480 x 240 x 160 cells: two steps per pass against single steps, 9 steps (pairs + one);
- 64 batched 2D simulations of 256..1024 cells a side (30.2 M cells): K = 8 steps per pass against one
  step per launch, 52 steps (K-step launches + pairs for the rest)."""
import numpy as np
import pytest

import beamz_amd as bz
from beamz_amd.simulation.batch import FDTDBatch
from beamz_amd.workloads import batch_2d, soi_mmi_3d

pytestmark = pytest.mark.gpu

COMP3 = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def test_full_size_3d_two_step_pass_equals_single_steps(hip):
    torch = hip
    out = {}
    for tb2 in (False, True):
        mesh, d3, t, mons = soi_mmi_3d(480, 240, 160, steps=9)
        sim = bz.FDTD(mesh, time=t)
        sim.grid.tb2 = tb2
        sim.run(live=False, save=False)
        r = {k: sim.field(k).clone() for k in COMP3}
        for q, m in enumerate(mons):
            for k, v in m.dft.items():
                r[f"dft{q}{k}"] = np.array(v)
        out[tb2] = r
        del sim
        torch.cuda.empty_cache()
    for k in COMP3:
        assert torch.isfinite(out[True][k]).all(), k
        assert torch.equal(out[True][k], out[False][k]), k
    for k in out[False]:
        if k.startswith("dft"):
            assert np.array_equal(out[True][k], out[False][k]), k
    assert any(float(out[True][k].abs().max()) > 0 for k in COMP3)


def test_full_size_batch_k_steps_equal_one_step(hip):
    torch = hip
    res = {}
    for key, kw in (("one", dict(kstep=0, pairs=False)), ("default", {})):
        members = batch_2d(nsims=64, steps=52)
        sims = [bz.FDTD(m, time=t) for m, t, _ in members]
        b = FDTDBatch(sims, **kw)
        b.initialize()
        if key == "default":
            assert b.engine.kstep == 8 and b.cells == 30_243_369
        b.advance(52)
        for s in sims:
            s.finalize_simulation()
        res[key] = ([{k: s.field(k).clone() for k in ("Ez", "Hx", "Hy")} for s in sims],
                    [{k: np.array(v) for k, v in mon.dft.items()} for _, _, ms in members for mon in ms])
        del b, sims
        torch.cuda.empty_cache()
    for fa, fb in zip(res["one"][0], res["default"][0]):
        for k in fa:
            assert torch.isfinite(fb[k]).all(), k
            assert torch.equal(fa[k], fb[k]), k
    for da, db in zip(res["one"][1], res["default"][1]):
        for k in da:
            assert np.array_equal(da[k], db[k]), k
