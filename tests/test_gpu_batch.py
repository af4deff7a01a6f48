"""GPU: the batched 2D engine (many uneven simulations per launch).

A member of a batch must take bit-for-bit the steps it takes alone, whatever
the tile height; graph replays must equal eager launches after every replay;
and a member must match the FP64 oracle within 1e-4 rel-L2 (fields, DFT)."""
import numpy as np
import pytest

import beamz_amd as bz
from beamz_amd.simulation.batch import FDTDBatch
from beamz_amd.workloads import batch_member_2d
from oracle import fdtd_ref as R

pytestmark = pytest.mark.gpu

SIDES = (96, 131, 170)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128); b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def members(steps=300, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for n in SIDES:
        mesh, t, mons = batch_member_2d(n, rng, steps=steps, design_um=1.0)
        out.append((bz.FDTD(mesh, time=t), mons))
    return out


def state(sim, mons):
    d = {k: np.real(getattr(sim, k)).copy() for k in ("Ez", "Hx", "Hy")}
    for m in mons:
        for k, v in m.dft.items():
            d[m.name + k] = np.array(v)
    return d


def test_batch_equals_single_runs_and_oracle(hip):
    single = []
    for sim, mons in members():
        sim.grid.kstep2d = 0                 # alone: one step per launch; in the batch: K steps per pass
        sim.run(live=False, save=False)
        single.append(state(sim, mons))
    mem = members()
    batch = FDTDBatch([s for s, _ in mem], seg=8, chunk=64)
    batch.run()
    for (sim, mons), ref in zip(mem, single):
        got = state(sim, mons)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (sim.nx, k)
    # one member against the FP64 oracle
    sim, mons = mem[1]
    ref = R.RefSim(sim.epsilon_r, sim.sigma, sim.dx, sim.dy, sim.dt, sim.num_steps, sources=sim.sources,
                   monitors=list(mons)).run()
    for k in ("Ez", "Hx", "Hy"):
        assert rel_l2(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, k
    for q, m in enumerate(mons):
        for k in ("Ez", "Hx", "Hy"):
            F = R.dft_of_series(np.array(ref.mon_series[q][k]), ref.mon_series[q]["t"], 2 * np.pi * m.dft_frequencies)
            assert rel_l2(m.dft[k], F) < 1e-4, (m.name, k)
    T = mons[1].dft_flux() / mons[0].dft_flux()
    assert T.shape == (8,) and np.all(np.isfinite(T))


def varied_members(seed=5):
    """Members that differ in everything but the polarisation: side, resolution (so time step), time axis
    length, DFT frequency count, record interval and whether series are kept."""
    rng = np.random.default_rng(seed)
    out = []
    for n, steps, dx, nf, every, keep in ((96, 300, 25e-9, 8, 1, True), (131, 221, 30e-9, 5, 3, False),
                                          (170, 262, 25e-9, 3, 2, True)):
        mesh, t, mons = batch_member_2d(n, rng, steps=steps, dx=dx, n_freq=nf, design_um=1.0)
        mons[1].record_interval = every
        mons[1].should_record_fields = keep
        out.append((bz.FDTD(mesh, time=t), mons))
    return out


def varied_state(sim, mons):
    d = state(sim, mons)
    for m in mons:
        for k, v in m.fields.items():
            if len(v):
                d[m.name + "series" + k] = np.array(v)
    return d


@pytest.mark.parametrize("kstep,pairs", [(None, None), (0, True), (0, False)])
def test_batch_members_with_their_own_time_axes_and_monitors(hip, kstep, pairs):
    """Members with different time steps, time axis lengths (a member that ends first stops there while the
    others go on), DFT frequency lists and record intervals take the bits they take alone, through the K-step,
    two-step and one-step kernels; one member against the FP64 oracle."""
    single = []
    for sim, mons in varied_members():
        sim.grid.kstep2d = 0
        sim.run(live=False, save=False)
        single.append(varied_state(sim, mons))
    mem = varied_members()
    batch = FDTDBatch([s for s, _ in mem], seg=8, chunk=50, kstep=kstep, pairs=pairs)
    batch.run()
    # point groups: member 0's two monitors share one; members 1 and 2 have their own time steps or frequency
    # counts, and their output monitors their own record intervals
    assert len(batch.engine.groups) == 5 and batch.engine.grp is not None
    assert (batch.engine.kstep > 0) == (kstep is None)
    assert sum(any("series" in k for k in ref) for ref in single) == 2     # members 0 and 2 keep series
    for (sim, mons), ref, ln in zip(mem, single, (300, 221, 262)):
        assert sim.current_step == ln
        got = varied_state(sim, mons)
        assert set(got) == set(ref), (sim.nx, sorted(set(got) ^ set(ref)))
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (sim.nx, k)
    sim, mons = mem[1]
    ref = R.RefSim(sim.epsilon_r, sim.sigma, sim.dx, sim.dy, sim.dt, sim.num_steps, sources=sim.sources,
                   monitors=list(mons)).run()
    for k in ("Ez", "Hx", "Hy"):
        assert rel_l2(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, k
    m = mons[0]                               # records every step, as the oracle does
    F = R.dft_of_series(np.array(ref.mon_series[0]["Ez"]), ref.mon_series[0]["t"], 2 * np.pi * m.dft_frequencies)
    assert rel_l2(m.dft["Ez"], F) < 1e-4


def test_batch_graph_replays_equal_eager(hip):
    torch = hip
    a = members(steps=320)
    b = members(steps=320)
    eager = FDTDBatch([s for s, _ in a], seg=8, graph=False, chunk=40, kstep=0)
    graph = FDTDBatch([s for s, _ in b], seg=8, graph=True, chunk=40)     # K-step launches, captured
    eager.initialize()
    graph.initialize()
    assert graph.engine.use_graph and graph.engine.kstep
    for r in range(8):                      # the 40-step graph is captured once and replayed 8 times
        eager.advance(40)
        graph.advance(40)
        junk = torch.randn(1 << 20, device="cuda").sum()       # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for (sa, _), (sb, _) in zip(a, b):
            for k in ("Ez", "Hx", "Hy"):
                assert torch.equal(sa.field(k), sb.field(k)), (r, sa.nx, k)
    for (sa, ma), (sb, mb) in zip(a, b):
        sa.finalize_simulation(); sb.finalize_simulation()
        A, B = state(sa, ma), state(sb, mb)
        for k in A:
            assert np.array_equal(A[k], B[k]), k


@pytest.mark.parametrize("kstep", [0, None])
def test_batch_autotune_keeps_results(hip, kstep):
    ref_mem = members(steps=200)
    FDTDBatch([s for s, _ in ref_mem], seg=8, kstep=0).run()
    mem = members(steps=200)
    batch = FDTDBatch([s for s, _ in mem], seg=8, kstep=kstep)
    best, timings = batch.autotune(steps=10)
    if kstep is None:
        segs = batch.engine.KCANDIDATES
        assert (batch.engine.kstep, batch.engine.krows, batch.engine.kwaves) == best
    else:
        segs = batch.engine.SEGS2 if batch.engine.pairs else FDTDBatch.SEGS
    assert best in segs and set(timings) == set(segs)
    batch.advance(batch.num_steps)
    for (s, m), (r, rm) in zip(mem, ref_mem):
        s.finalize_simulation()
        A, B = state(s, m), state(r, rm)
        for k in A:
            assert np.array_equal(A[k], B[k]), (best, k)


@pytest.mark.parametrize("seg2", [2, 7, 64])
def test_two_step_pairs_equal_single_steps(hip, seg2):
    """Two steps per launch (staged intermediate monitor samples) must give the bits of one step per launch,
    with an odd step count (pairs plus a last single step) and any pair-tile height."""
    single = members(steps=301)
    FDTDBatch([s for s, _ in single], pairs=False, kstep=0).run()
    pair = members(steps=301)
    b = FDTDBatch([s for s, _ in pair], pairs=True, seg2=seg2, chunk=75, kstep=0)
    b.run()
    assert b.engine.pairs
    for (s, m), (r, rm) in zip(pair, single):
        A, B = state(s, m), state(r, rm)
        for k in A:
            assert np.array_equal(A[k], B[k]), (seg2, s.nx, k)


@pytest.mark.parametrize("count", [7, 10])
def test_pair_snapshots_equal_one_step_snapshots(hip, count):
    """Snapshot boxes written by two-step launches (slots 2q, 2q+1, and a last single step after the pairs)
    and by K-step launches (slots qK+k, then pairs and a single step for the rest of the run) equal those of
    one step per launch, bit for bit, as do the fields."""
    torch = hip
    from beamz_amd.simulation.engine import Fused2D
    got = {}
    for key, kstep, pairs in (("one", 0, False), ("pairs", 0, True), ("k3", 3, True), ("k4", 4, True)):
        sims = [s for s, _ in members(steps=64)]
        for s in sims:
            s.initialize_simulation(save=False, live=False)
        bufs = [torch.full((count, 3, 30, 40), -7.0, dtype=torch.float32, device="cuda") for _ in sims]
        snaps = [(10, 12, 40, 30, b) for b in bufs]
        eng = Fused2D([s.grid for s in sims], seg=8, monitors=False, snapshots=snaps, pairs=pairs, seg2=7,
                      kstep=kstep, krows=8, kwaves=4)
        for s in sims:
            s.grid._f2d = False
        eng.run(0, 5)
        eng.run(5, count)
        assert eng.pairs == pairs and eng.kstep == kstep
        got[key] = ([b.cpu().numpy() for b in bufs], [{k: s.field(k).cpu().numpy() for k in ("Ez", "Hx", "Hy")}
                                                       for s in sims])
    for key in ("pairs", "k3", "k4"):
        for a, b in zip(got["one"][0], got[key][0]):
            assert not np.any(a == -7.0)
            assert np.array_equal(a, b), key
        for fa, fb in zip(got["one"][1], got[key][1]):
            for k in fa:
                assert np.array_equal(fa[k], fb[k]), (key, k)


@pytest.mark.parametrize("K,rows,waves,pairs", [(1, 8, 4, False), (3, 16, 4, False), (8, 8, 8, True),
                                                (15, 8, 4, False), (31, 16, 4, False), (6, 4, 8, False),
                                                (12, 4, 16, False), (31, 8, 16, False)])
def test_k_steps_per_pass_equal_single_steps(hip, K, rows, waves, pairs):
    """K steps per pass (register bands, LDS edge exchange, staged monitor samples) must give the bits of one
    step per launch: 301 steps in runs of 75 (K-step launches, then pairs or single steps for the rest), inner
    tiles from 62 x 62 down to 2 x 2 cells, uneven members, sources and DFT monitors."""
    single = members(steps=301)
    FDTDBatch([s for s, _ in single], pairs=False, kstep=0).run()
    mem = members(steps=301)
    b = FDTDBatch([s for s, _ in mem], pairs=pairs, kstep=K, krows=rows, kwaves=waves, chunk=75)
    b.run()
    assert b.engine.kstep == K and b.engine.pairs == pairs
    for (s, m), (r, rm) in zip(mem, single):
        A, B = state(s, m), state(r, rm)
        for k in A:
            assert np.array_equal(A[k], B[k]), (K, rows, waves, s.nx, k)


def test_new_materials_after_reset_take_effect(hip):
    """The K-step kernel's per-cell coefficients follow the materials: writing a member's permittivity between
    runs (as a design loop does) gives the bits of a batch built with that permittivity from the start."""
    torch = hip
    mem = members(steps=200)
    b = FDTDBatch([s for s, _ in mem], chunk=48)
    b.run()
    assert b.engine.kstep
    s0 = mem[0][0]
    eps_new = s0.grid.eps.clone()
    eps_new[20:40, 20:40] = 3.476 ** 2
    b.reset()
    s0.grid.eps.copy_(eps_new)
    b.advance(b.num_steps)
    for s, _ in mem:
        s.finalize_simulation()
    fresh = members(steps=200)
    f = FDTDBatch([s for s, _ in fresh], chunk=48)
    f.initialize()
    fresh[0][0].grid.eps.copy_(eps_new)
    f.advance(f.num_steps)
    for s, _ in fresh:
        s.finalize_simulation()
    for (s, m), (r, rm) in zip(mem, fresh):
        A, B = state(s, m), state(r, rm)
        for k in A:
            assert np.array_equal(A[k], B[k]), (s.nx, k)
    # and the change mattered: the member differs from a run on the original permittivity
    orig = members(steps=200)
    FDTDBatch([s for s, _ in orig], chunk=48).run()
    assert not np.array_equal(np.real(orig[0][0].Ez), np.real(mem[0][0].Ez))
