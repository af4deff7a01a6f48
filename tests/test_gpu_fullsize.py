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


def test_c1_grid_matches_fp64_oracle(hip):
    """The reference's own CPU case at its full grid: 800 x 400 cells at lambda/20, the 0.5 um Si guide, the
    20-cell PML, the point pulse at x = 50 and a 16-wavelength DFT line, for the first 600 of its 4000 steps
    (the FP64 oracle takes seconds for that). The line sits at x = 100 instead of 700, where the pulse arrives
    within those steps. Fields, the line's series and its DFT within 1e-4 rel-L2 of the oracle, through the
    default K-step path."""
    from beamz_amd.workloads import waveguide_2d
    from oracle import fdtd_ref as R
    d, dx, t, mon = waveguide_2d(800, 400, steps=600, mon_x=100)
    sim = bz.FDTD(d, time=t, resolution=dx)
    sim.run(live=False, save=False)
    assert sim.grid._f2d.kstep
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, float(t[1] - t[0]), len(t), sources=d.sources,
                   monitors=[mon]).run()

    def rel(a, b):
        a, b = np.asarray(a, np.complex128), np.asarray(b, np.complex128)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    assert np.abs(np.array(ref.mon_series[0]["Ez"])).max() > 1e-3 * np.abs(ref.F["Ez"]).max()   # the pulse arrived
    for k in ("Ez", "Hx", "Hy"):
        assert rel(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, k
        assert rel(np.array(mon.fields[k]), np.array(ref.mon_series[0][k])) < 1e-4, k
        F = R.dft_of_series(np.array(ref.mon_series[0][k]), ref.mon_series[0]["t"], 2 * np.pi * mon.dft_frequencies)
        assert rel(mon.dft[k], F) < 1e-4, k
