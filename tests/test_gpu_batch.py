"""GPU: batches of RVEs of uneven size, geometry, material law and coupling.

Each RVE of a batch follows micropp<3>'s per-GP algorithm with the
single-RVE kernels' arithmetic (same bodies, same reduction order), so the
bar against a micropp<3> instance of that one RVE is bit-for-bit equality of
sigma, ctan and the cost / convergence flags, step after step.  Against the
oracle (CPU restatement of the reference) the bar is 1e-6 relative.
"""
import numpy as np
import pytest

from oracle import bindings as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


SPECS = {
    "elastic": [(0, 1e6, 0.3), (0, 1e7, 0.25), (0, 1e7, 0.25)],
    "plastic": [(1, 1e6, 0.3, 1e5, 1e3), (0, 1e7, 0.25), (1, 2e6, 0.3, 1e4, 5e2)],
    "damage": [(2, 1e6, 0.3, 0, 0, 1e5), (0, 1e7, 0.25), (2, 1e6, 0.3, 0, 0, 1e5)],
    # the J2 / fibre case of the batched-plasticity workload: Sy = 1e4, Ka = 1e5
    "j2": [(1, 1e6, 0.3, 1e5, 1e4), (0, 1e7, 0.25), (0, 1e7, 0.25)],
}
AMP = {"elastic": 2e-3, "plastic": 2e-3, "damage": 0.02, "j2": 2e-2}


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    return 0.0 if not m.any() else np.abs(a[m] - b[m]).max() / max(np.abs(b[m]).max(), 1e-300)


def mixed_batch(mp):
    """Uneven sizes (two RVEs share one), five geometries, three laws, three couplings."""
    return [
        dict(size=5, micro_type=mp.MIC_SPHERE, geo=(0.3, 0, 0, 0), kind="plastic", coupling=mp.FE_FULL),
        dict(size=6, micro_type=mp.MIC_LAYER_Y, geo=(0.4, 0, 0, 0), kind="damage", coupling=mp.FE_ONE_WAY),
        dict(size=7, micro_type=mp.MIC_CILI_FIB_Z, geo=(0.25, 0, 0, 0), kind="elastic", coupling=mp.FE_FULL),
        dict(size=5, micro_type=mp.MIC3D_RANDOM_SPHERES, geo=(7.0, 4.0, 0.1, 0.25), kind="plastic",
             coupling=mp.FE_ONE_WAY),
        dict(size=(5, 6, 7), micro_type=mp.MIC_HOMOGENEOUS, geo=(0.1, 0.1, 0.1, 0.1), kind="elastic",
             coupling=mp.FE_LINEAR),
        dict(size=6, micro_type=mp.MIC_CILI_FIB_Z, geo=(0.2, 0, 0, 0), kind="j2", coupling=mp.FE_FULL),
    ]


def strains_for(desc, steps, seed):
    rng = np.random.default_rng(seed)
    amp = AMP[desc["kind"]]
    d = rng.uniform(-amp, amp, 6)
    return [d * (t + 1) / steps for t in range(steps)]


KW = dict(lin_stress=0, subiterations=1, nsubiterations=3, nr_max_its=4, cg_rel_tol=1e-12, cg_max_its=3000)


def run_batch(mp, descs, steps, graph_chunk=0, **kw):
    rves = [dict(size=d["size"], micro_type=d["micro_type"], geo=d["geo"],
                 materials=[mp.material(*s) for s in SPECS[d["kind"]]], coupling=d["coupling"]) for d in descs]
    b = mp.Batch(rves, graph_chunk=graph_chunk, **kw)
    ramps = [strains_for(d, steps, 100 + i) for i, d in enumerate(descs)]
    rec = []
    for t in range(steps):
        b.set_strains(np.array([r[t] for r in ramps]))
        b.homogenize()
        rec.append((b.stress(), b.ctan(), b.info()))
        b.update_vars()
    return rec, b.stats()


def run_single(lib_micro, descs, steps):
    out = []
    for i, d in enumerate(descs):
        m = lib_micro(d)
        ramp = strains_for(d, steps, 100 + i)
        rec = []
        for t in range(steps):
            m.set_strain(0, ramp[t])
            m.homogenize()
            rec.append((m.get_stress(0), m.get_ctan(0), (m.is_non_linear(0), m.get_cost(0), int(m.has_converged(0)),
                                                        int(m.has_subiterated(0)))))
            m.update_vars()
        out.append(rec)
    return out


@pytest.mark.parametrize("its_with_A0", [0, 2])
def test_batch_equals_single_instances(mp, its_with_A0):
    """its_with_A0 > 0: Newton on the initial matrix (use_A0) in the batch and in every instance; a seventh
    RVE repeats the first one's descriptor (another coupling), so the two share one initial matrix."""
    descs = mixed_batch(mp)
    kw = dict(KW)
    if its_with_A0:
        descs.append(dict(descs[0], coupling=mp.FE_ONE_WAY))
        kw.update(use_A0=1, its_with_A0=its_with_A0)
    steps = 3
    rec, stats = run_batch(mp, descs, steps, **kw)
    print("batch stats (CG its queued, graph chunk, graph launches, ms):", stats, flush=True)

    def gpu_micro(d):
        return mp.Micropp3(d["size"], d["micro_type"], d["geo"], [mp.material(*s) for s in SPECS[d["kind"]]],
                           ngp=1, coupling=[d["coupling"]], **kw)

    single = run_single(gpu_micro, descs, steps)
    for t in range(steps):
        sig, ctan, info = rec[t]
        for i in range(len(descs)):
            s1, c1, f1 = single[i][t]
            print(f"step {t} rve {i}: info {tuple(info[i])} vs {f1}; sigma rel {rel(sig[i], s1):.2e}", flush=True)
            assert np.array_equal(sig[i], s1)
            assert np.array_equal(ctan[i], c1)
            assert tuple(int(v) for v in info[i]) == tuple(int(v) for v in f1)


def test_batch_matches_oracle(mp, oracle):
    descs = mixed_batch(mp)
    steps = 3
    rec, _ = run_batch(mp, descs, steps, **KW)

    def orc_micro(d):
        p, k = ob.make_params(d["size"], d["micro_type"], d["geo"], [ob.material(*s) for s in SPECS[d["kind"]]],
                              ngp=1, coupling=[d["coupling"]], **KW)
        return oracle.micro(p, k)

    single = run_single(orc_micro, descs, steps)
    for t in range(steps):
        sig, ctan, info = rec[t]
        for i in range(len(descs)):
            s1, c1, f1 = single[i][t]
            print(f"step {t} rve {i}: sigma rel {rel(sig[i], s1):.2e} ctan rel {rel(ctan[i], c1):.2e}", flush=True)
            assert rel(sig[i], s1) < 1e-6
            assert rel(ctan[i], c1) < 1e-6
            assert (int(info[i][0]), int(info[i][2]), int(info[i][3])) == (int(f1[0]), int(f1[2]), int(f1[3]))


@pytest.mark.parametrize("chunk", [2, 16])
def test_graph_chunk_does_not_change_results(mp, chunk):
    descs = mixed_batch(mp)[:3]
    a, _ = run_batch(mp, descs, 2, graph_chunk=chunk, **KW)
    b, st = run_batch(mp, descs, 2, graph_chunk=0, **KW)
    assert st[1] in (2, 4, 8, 16, 32)
    for (sa, ca, ia), (sb, cb, ib) in zip(a, b):
        assert np.array_equal(sa, sb) and np.array_equal(ca, cb) and np.array_equal(ia, ib)


def test_batch_damage_layers_uneven(mp, oracle):
    """The batched-damage workload at small sizes: fibre layers of seeded width, damage matrix, elastic fibre,
    uneven sizes, a seeded strain ramp per RVE, Newton tolerance 1e-6."""
    rng = np.random.default_rng(2)
    sizes = [5, 7, 6, 9, 5, 8]
    descs = [dict(size=s, micro_type=mp.MIC_LAYER_Y, geo=(rng.uniform(0.1, 0.3), 0, 0, 0), kind="damage",
                  coupling=mp.FE_ONE_WAY) for s in sizes]
    kw = dict(lin_stress=0, nr_max_its=10, nr_rel_tol=1e-6, cg_rel_tol=1e-12, cg_max_its=3000)
    steps = 4
    rvs = [dict(size=d["size"], micro_type=d["micro_type"], geo=d["geo"],
                materials=[mp.material(*s) for s in SPECS["damage"]], coupling=d["coupling"]) for d in descs]
    b = mp.Batch(rvs, **kw)
    ramps = []
    for i in range(len(descs)):
        d = rng.normal(size=6)
        ramps.append([d / np.linalg.norm(d) * 0.04 * (t + 1) / steps for t in range(steps)])
    orcs = []
    for d in descs:
        p, k = ob.make_params(d["size"], d["micro_type"], d["geo"], [ob.material(*s) for s in SPECS["damage"]],
                              ngp=1, coupling=[d["coupling"]], **kw)
        orcs.append(oracle.micro(p, k))
    for t in range(steps):
        b.set_strains(np.array([r[t] for r in ramps]))
        b.homogenize()
        sig = b.stress()
        for i, o in enumerate(orcs):
            o.set_strain(0, ramps[i][t])
            o.homogenize()
            print(f"step {t} rve {i}: sigma rel {rel(sig[i], o.get_stress(0)):.2e}", flush=True)
            assert rel(sig[i], o.get_stress(0)) < 1e-6
            o.update_vars()
        b.update_vars()


def test_batch_rejects_bad_input(mp):
    with pytest.raises(mp.MppdError):
        mp.Batch([dict(size=5, micro_type=mp.MIC_SPHERE, geo=(0.2, 0, 0, 0),
                       materials=[mp.material(0, 1e6, 0.3)] * 3, coupling=mp.MIX_RULE_CHAMIS)])
    with pytest.raises(mp.MppdError):
        mp.Batch([dict(size=5, micro_type=99, geo=(0.2, 0, 0, 0), materials=[mp.material(0, 1e6, 0.3)] * 3)])
    with pytest.raises(mp.MppdError):
        mp.Batch([dict(size=5, micro_type=1, geo=(0.2, 0, 0, 0), materials=[mp.material(0, 1e6, 0.3)] * 3)],
                 graph_chunk=3)
    with pytest.raises(mp.MppdError):
        mp.Batch([dict(size=5, micro_type=1, geo=(0.2, 0, 0, 0), materials=[mp.material(0, 1e6, 0.3)] * 3)],
                 spmv_chunk=10)


def test_batch_newton_stops_unconverged_at_the_cap(mp):
    """The reference's Newton tests the residual only before an iteration it will run: at nr_max_its it reports
    non-convergence even when the last residual passes (src/solve.cpp:43-78).  A yielding J2 RVE whose third
    residual passes 1e-6 must come out unconverged and subiterated, as a micropp<3> instance reports it."""
    mats = [(1, 1e6, 0.3, 1e5, 2e3), (0, 1e7, 0.25), (0, 1e7, 0.25)]
    kw = dict(lin_stress=0, subiterations=1, nsubiterations=3, nr_max_its=3, nr_rel_tol=1e-6, cg_rel_tol=1e-10,
              cg_max_its=3000)
    eps = np.random.default_rng(7).uniform(-1, 1, (4, 6))[3] * 6e-3 / 3
    rv = dict(size=6, micro_type=mp.MIC_SPHERE, geo=(0.3, 0, 0, 0), materials=[mp.material(*s) for s in mats],
              coupling=mp.FE_FULL)
    b = mp.Batch([rv, dict(rv, coupling=mp.FE_ONE_WAY)], **kw)
    b.set_strains(np.array([eps, eps]))
    b.homogenize()
    m = mp.Micropp3(6, mp.MIC_SPHERE, (0.3, 0, 0, 0), rv["materials"], ngp=1, coupling=[mp.FE_FULL], **kw)
    m.set_strain(0, eps)
    m.homogenize()
    info = b.info()
    print("batch info", info.tolist(), "instance", m.get_cost(0), m.has_converged(0), m.has_subiterated(0), flush=True)
    assert np.array_equal(b.stress()[0], m.get_stress(0))
    assert (int(info[0][1]), int(info[0][2]), int(info[0][3])) == (m.get_cost(0), int(m.has_converged(0)),
                                                                   int(m.has_subiterated(0)))
    assert (int(info[0][2]), int(info[0][3])) == (0, 1)  # the case: subiterated, last Newton stopped at the cap


@pytest.mark.parametrize("chunk,order", [(9, 0), (81, 1), (27, 0)])
def test_spmv_shape_does_not_change_results(mp, chunk, order):
    """The autotuned SpMV shapes (ELL slots per load round, block order) change only load scheduling: a batch
    with a fixed shape gives bit-identical results to the tuned one."""
    descs = mixed_batch(mp)[:4]
    a, _ = run_batch(mp, descs, 2, spmv_chunk=chunk, block_order=order, **KW)
    rves = [dict(size=d["size"], micro_type=d["micro_type"], geo=d["geo"],
                 materials=[mp.material(*s) for s in SPECS[d["kind"]]], coupling=d["coupling"]) for d in descs]
    tuned = mp.Batch(rves, **KW)
    g, c, x = tuned.tuning()
    print(f"tuned: graph chunk {g}, SpMV slots {c}, XCD order {x}", flush=True)
    assert g in (2, 4, 8, 16, 32) and c in (9, 27, 81) and x in (0, 1)
    b, _ = run_batch(mp, descs, 2, **KW)
    for (sa, ca, ia), (sb, cb, ib) in zip(a, b):
        assert np.array_equal(sa, sb) and np.array_equal(ca, cb) and np.array_equal(ia, ib)


@pytest.mark.parametrize("its_with_A0", [0, 2])
def test_device_newton_equals_host_newton(mp, monkeypatch, its_with_A0):
    """The FP64 batch's Newton loop decides on the device (per-RVE state, a captured head graph, CG chunk graphs
    and a captured finishing graph per Newton iteration, one host read of the states after it).  Against the
    host-driven loop of the same batch (MPPD_BATCH_HOST_NEWTON=1, forced from the host): sigma, ctan, cost and
    flags bit for bit, with fewer host synchronisations (at most about one per batched Newton iteration plus a
    fixed few per solve)."""
    descs = mixed_batch(mp)
    kw = dict(KW, use_A0=1, its_with_A0=its_with_A0) if its_with_A0 else dict(KW)
    steps = 2
    rec_d, st_d = run_batch(mp, descs, steps, **kw)
    monkeypatch.setenv("MPPD_BATCH_HOST_NEWTON", "1")
    rec_h, st_h = run_batch(mp, descs, steps, **kw)
    print(f"host syncs of the last homogenize: device loop {st_d[4]:.0f}, host loop {st_h[4]:.0f}; batched Newton "
          f"iterations {st_d[5]:.0f} / {st_h[5]:.0f}", flush=True)
    for t in range(steps):
        for x, y in zip(rec_d[t], rec_h[t]):
            assert np.array_equal(x, y), f"step {t}"
    assert st_d[4] < st_h[4]
    assert st_d[5] > 0
