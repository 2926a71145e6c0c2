"""Batches of independent scenes (mpm_batch_*): advancing handles through a
batch -- recorded into a HIP graph and replayed -- gives the same particles,
bit for bit, as advancing each handle on its own (the tiled step sums in a
fixed order, so both runs compute the same floats)."""
import numpy as np
import pytest

from conftest import state_errors
from mpm_amd import MATERIAL_SNOW, MATERIAL_WC_LIQUID, Batch, Engine, scenes

pytestmark = pytest.mark.gpu


def make(sc):
    st = sc.state()
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    return eng


def small_batch():
    return [
        scenes.rubber_elastic(3000, 61, N=32, velocity=(0.5, -1.0, 0.2)),
        scenes.small_scene(MATERIAL_SNOW, 5000, 62, N=48, velocity=(0.0, -2.0, 0.0)),
        scenes.small_scene(MATERIAL_WC_LIQUID, 4000, 63, N=32),
        scenes.small_scene(MATERIAL_SNOW, 2000, 64, N=16, center=(0.4, 0.6, 0.5), radius=0.1),
    ]


def run_alone(scs, k):
    out = []
    for sc in scs:
        e = make(sc)
        e.advance(k)
        d = e.download()
        d["t"] = e.time()
        out.append(d)
        e.close()
    return out


@pytest.mark.parametrize("graph,fuse,multi_sort", [(1, 1, 1), (1, 0, 1), (0, 1, 1), (1, 1, 0)])
def test_batch_equals_individual_runs(gpu_lib, graph, fuse, multi_sort):
    scs = small_batch()
    # 16 eager substeps, two recorded cycles of 32, 20 eager; with "fuse" the
    # eager substeps after the first (which primes each scene) run as
    # multi-scene launches too; without "multi_sort" each scene keeps its own
    # kernel chain (its sorts cannot join the multi-scene launches)
    k = 100
    ref = run_alone(scs, k)
    engines = [make(sc) for sc in scs]
    b = Batch(engines)
    b.set_option("graph", graph)
    b.set_option("fuse", fuse)
    b.set_option("multi_sort", multi_sort)
    b.advance(k)
    outs = [e.download() for e in engines]
    launches = b.get_option("graph_launches")
    fused = b.get_option("fused_captures")
    eager = b.get_option("fused_eager_substeps")
    b.close()
    fuse = fuse and multi_sort
    assert fused == (1 if graph and fuse else 0)
    if not fuse:
        assert eager == 0
    elif graph:
        assert 16 + 20 - 2 <= eager <= 16 + 20 - 1, eager
    else:
        assert k - 2 <= eager <= k - 1, eager
    for e, r in zip(engines, ref):
        assert e.time() == r["t"]
    for o, r in zip(outs, ref):
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(o[f], r[f]), f
    if graph:
        assert launches == 2
    else:
        assert launches == 0
    for e in engines:
        e.close()


@pytest.mark.parametrize("chains", [1, 2, 3])
def test_batch_chains_equal_individual_runs(gpu_lib, chains):
    """The scenes split into 1-3 chains of launches on streams of their own
    (the default, 4, is the other tests'): each scene's particles as alone."""
    scs = small_batch()
    k = 100
    ref = run_alone(scs, k)
    engines = [make(sc) for sc in scs]
    b = Batch(engines)
    b.set_option("chains", chains)
    assert b.get_option("chains") == chains
    b.advance(k)
    outs = [e.download() for e in engines]
    assert b.get_option("fused_captures") == 1
    b.close()
    for e, o, r in zip(engines, outs, ref):
        assert e.time() == r["t"]
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(o[f], r[f]), f
        e.close()


def test_batch_recaptures_after_a_change(gpu_lib):
    scs = small_batch()[:2]
    engines = [make(sc) for sc in scs]
    b = Batch(engines)
    b.advance(48)          # eager 16, one cycle
    engines[1].set_option("chunk", 128)   # changes the scene: the next cycle is recorded again
    b.advance(64)
    outs = [e.download() for e in engines]
    times = [e.time() for e in engines]
    b.close()
    # the same sequence without a batch
    ref = []
    for j, sc in enumerate(scs):
        e = make(sc)
        e.advance(48)
        if j == 1:
            e.set_option("chunk", 128)
        e.advance(64)
        ref.append((e.download(), e.time()))
        e.close()
    for o, t, (r, rt) in zip(outs, times, ref):
        assert t == rt
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(o[f], r[f]), f
    for e in engines:
        e.close()


def test_batched_workload_runs(gpu_lib):
    # the 64-scene workload at 1/20 of its particle counts, a few cycles
    scs = scenes.batched_scenes(scale=0.05)
    engines = [make(sc) for sc in scs]
    b = Batch(engines)
    b.advance(80)
    for e in engines:
        out = e.download()
        assert np.isfinite(out["x"]).all()
    b.close()
    for e in engines:
        e.close()


def test_batch_of_scenes_with_other_radix_digits(gpu_lib):
    """A scene that sorts with 6- or 4-bit radix digits (an autotuner choice)
    joins a batch, whose sort uses 8-bit digits: the stable sort gives the same
    order, so the particles are the same bits as the scene's own run."""
    scs = small_batch()[:3]
    bits = (6, 4, 8)
    k = 60

    def with_bits(sc, b):
        e = make(sc)
        e.set_option("radix_bits", b)
        return e

    ref = []
    for sc, b in zip(scs, bits):
        e = with_bits(sc, b)
        e.advance(k)
        ref.append(e.download())
        e.close()
    engines = [with_bits(sc, b) for sc, b in zip(scs, bits)]
    bt = Batch(engines)
    bt.advance(k)
    outs = [e.download() for e in engines]
    bt.close()
    for o, r in zip(outs, ref):
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(o[f], r[f]), f
    for e in engines:
        e.close()


@pytest.mark.parametrize("share", [1, 20])
def test_batch_workgroups_taking_further_items(gpu_lib, share):
    """Each scene's share of the fused pass is sized from an estimate of its
    work items, and when a scene has more items than workgroups, every
    workgroup takes further items of it in turn. Shrinking the shares (option
    item_share: 1% leaves one workgroup per scene, 20% a few) sends the
    multi-scene passes through that path: the batch still equals every scene's
    own run bit for bit."""
    scs = small_batch()
    k = 100
    ref = run_alone(scs, k)
    engines = [make(sc) for sc in scs]
    b = Batch(engines)
    b.set_option("item_share", share)
    assert b.get_option("item_share") == share
    b.advance(k)
    outs = [e.download() for e in engines]
    assert b.get_option("fused_captures") == 1
    b.close()
    for o, r in zip(outs, ref):
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(o[f], r[f]), f
    for e in engines:
        e.close()


def test_full_batch_equals_individual_runs(gpu_lib):
    """The batched workload at full size (64 scenes, 3.35 M particles, N in
    {32, 48, 64}) for 48 substeps -- 16 per scene, then one recorded cycle of
    32 as multi-scene launches -- equals every scene's own run bit for bit."""
    scs = scenes.batched_scenes()
    assert len(scs) == 64
    k = 48
    engines = [make(sc) for sc in scs]
    b = Batch(engines)
    b.advance(k)
    outs = [e.download() for e in engines]
    fused = b.get_option("fused_captures")
    b.close()
    for e in engines:
        e.close()
    assert fused == 1
    for j, sc in enumerate(scs):
        e = make(sc)
        e.advance(k)
        r = e.download()
        e.close()
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(outs[j][f], r[f]), (sc.name, f)


def pick_oracle_scenes(scs, count=8):
    """Scenes covering every material kind and every grid size, smallest
    first within each (kind, N) pair, then the smallest others."""
    by = {}
    for j, sc in enumerate(scs):
        kind = int(sc.name.split("_k")[1].split("_")[0])
        by.setdefault((kind, sc.params.N), []).append(j)
    chosen = [min(js, key=lambda j: scs[j].n) for js in by.values()]
    rest = sorted((j for j in range(len(scs)) if j not in chosen), key=lambda j: scs[j].n)
    return sorted(chosen + rest[:max(0, count - len(chosen))])


def test_full_batch_against_oracle(gpu_lib):
    """The multi-scene kernels (grid update, cell sort, fused G2P/P2G of every
    scene in one launch each) against the oracle: the 64-scene workload with
    a sort every 2 substeps runs 2 substeps per scene, then one recorded
    cycle of 4 substeps (two multi-scene sorts); 8 scenes covering elastic,
    snow and liquid at N = 32, 48 and 64 are checked against the FP32 and FP64
    oracle after those 6 substeps (elastic: 1e-4; snow: 1e-3 and liquid: 1e-4,
    or the FP32 oracle's own distance from FP64)."""
    import oracle
    from test_gpu_parity import compare_with_oracle

    scs = scenes.batched_scenes()
    k = 6
    engines = []
    for sc in scs:
        e = make(sc)
        e.set_option("sort_every", 2)
        engines.append(e)
    b = Batch(engines)
    b.advance(k)
    fused = b.get_option("fused_captures")
    launches = b.get_option("graph_launches")
    outs = [e.download() for e in engines]
    b.close()
    for e in engines:
        e.close()
    assert fused == 1 and launches == 1
    picked = pick_oracle_scenes(scs)
    kinds = {int(scs[j].name.split("_k")[1].split("_")[0]) for j in picked}
    Ns = {scs[j].params.N for j in picked}
    assert kinds == {0, 1, 2} and Ns == {32, 48, 64}, (kinds, Ns)
    threads = oracle.check_threads()
    for j in picked:
        sc = scs[j]
        kind = int(sc.name.split("_k")[1].split("_")[0])
        tol = {0: 1e-4, 1: 1e-3, 2: 1e-4}[kind]
        compare_with_oracle(sc, sc.state(), outs[j], k, tol, fp32_noise=kind != 0, threads=threads,
                            label=f"{sc.name} ({sc.n} particles) in the batch")


def test_batch_with_different_sort_intervals(gpu_lib):
    """Scenes sorting every 16, 24 and 8 substeps share one recorded cycle (the
    least common multiple of their two-interval cycles, 96 substeps) of
    multi-scene launches; each equals its own run bit for bit."""
    scs = small_batch()[:3]
    every = (16, 24, 8)
    k = 48 + 96 + 20  # to the first common cycle start, one cycle, 20 more

    def with_se(sc, se):
        e = make(sc)
        e.set_option("sort_every", se)
        return e

    ref = []
    for sc, se in zip(scs, every):
        e = with_se(sc, se)
        e.advance(k)
        ref.append(e.download())
        e.close()
    engines = [with_se(sc, se) for sc, se in zip(scs, every)]
    b = Batch(engines)
    b.advance(k)
    outs = [e.download() for e in engines]
    assert b.get_option("fused_captures") == 1 and b.get_option("graph_launches") == 1
    b.close()
    for o, r in zip(outs, ref):
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(o[f], r[f]), f
    for e in engines:
        e.close()


def test_batch_autotune_restores_and_keeps_bits(gpu_lib):
    """The batch tuner times its candidates from a snapshot of every scene and
    puts the scenes back (same particles, time 0); then the batch under the
    chosen settings equals each scene run alone with the same settings: bit
    for bit while every particle stays in its tile between sorts; a sort
    interval long enough for a fast scene's particles to leave their tiles
    sends them to the direct path, whose global FP32 atomics add in hardware
    order (two runs alone differ as well), and that scene then agrees within
    the step tolerance."""
    scs = small_batch()
    engines = [make(sc) for sc in scs]
    b = Batch(engines)
    choice = b.autotune(trial_substeps=64)
    print("batch tuner:", choice)
    assert choice["choice"] == "own" or set(choice["choice"]) == {"sort_every", "chunk"}
    assert len(choice["trials"]) >= 2
    for e, sc in zip(engines, scs):
        assert e.time() == 0.0
        st = sc.state()
        d = e.download()
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(d[f], st[f].astype(np.float32)), f
    opts = [{k: e.get_option(k) for k in ("sort_every", "chunk", "radix_bits")} for e in engines]
    k = 150
    b.advance(k)
    outs = [e.download() for e in engines]
    b.close()
    for e in engines:
        e.close()
    for sc, o, out in zip(scs, opts, outs):
        e = make(sc)
        for name, v in o.items():
            e.set_option(name, v)
        e.advance(k)
        r = e.download()
        e.close()
        if all(np.array_equal(out[f], r[f]) for f in ("x", "v", "C", "F", "Jp")):
            continue
        errs = state_errors(out, r, sc.params.N)
        assert all(e < 1e-4 for e in errs.values()), (sc.name, o, errs)
