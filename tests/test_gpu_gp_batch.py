"""GPU: micropp<3>::homogenize running its FE GPs as a batch of identical RVE slots.

The reference's homogenize loops over the GPs of an instance (OpenMP,
src/homogenize.cpp:68-99).  Here the FE GPs of an instance run together as
working slots of one batch, in groups when the slots are fewer than the GPs.
Every slot runs the one-GP kernels' arithmetic, so the bar against the
one-GP-at-a-time path (MPPD_GP_BATCH=0) is bit-for-bit equality of sigma,
ctan, cost and flags, and of the device state u_n, u_k, vars_n, vars_k, step
after step -- across groups, idle slots, couplings, subiterations, history
allocation, update_vars and a restart.
"""
import numpy as np
import pytest

from oracle import bindings as ob

pytestmark = pytest.mark.gpu

J2 = [(1, 1e6, 0.3, 1e5, 2e3), (0, 1e7, 0.25), (0, 1e7, 0.25)]


@pytest.fixture(scope="module")
def mp():
    import micropp_amd

    micropp_amd.load()
    if not micropp_amd.gpu_available():
        pytest.skip("no HIP device")
    return micropp_amd


def couplings(mp):
    # 5 FE GPs: with 2 slots they run as groups of 2, 2 and 1 (one idle slot)
    return [mp.FE_FULL, mp.FE_ONE_WAY, mp.FE_LINEAR, mp.FE_FULL, mp.MIX_RULE_CHAMIS, mp.FE_ONE_WAY, mp.FE_FULL]


def ramps(ngp, steps, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    amps = [3e-3, 2e-4, 1e-3, 6e-3, 1e-3, 3e-3, 2e-4]  # FE GPs 0, 3, 5 yield; 1, 6 stay elastic
    for g in range(ngp):
        d = rng.uniform(-1, 1, 6) * amps[g % len(amps)]
        out.append([d * (t + 1) / steps for t in range(steps)])
    return out


def make(mp, **kw):
    cp = couplings(mp)
    return mp.Micropp3(6, mp.MIC_SPHERE, (0.3, 0, 0, 0), [mp.material(*s) for s in J2], ngp=len(cp), coupling=cp,
                       **kw)


def record(m):
    rec = []
    for g in range(m.ngp):
        st = tuple(m.state(g, w) for w in range(4))
        rec.append((m.get_stress(g), m.get_ctan(g), (m.is_non_linear(g), m.get_cost(g), int(m.has_converged(g)),
                                                     int(m.has_subiterated(g))), st))
    return rec


def step(m, rp, t, monkeypatch=None, slots=None):
    """One load step; the slot count (MPPD_GP_BATCH) is read at an instance's first homogenize."""
    if monkeypatch is not None:
        monkeypatch.setenv("MPPD_GP_BATCH", str(slots))
    for g in range(m.ngp):
        m.set_strain(g, rp[g][t])
    m.homogenize()
    r = record(m)
    m.update_vars()
    return r


def same(a, b, where):
    for g, (ra, rb) in enumerate(zip(a, b)):
        sa, ca, fa, xa = ra
        sb, cb, fb, xb = rb
        assert np.array_equal(sa, sb), f"{where} gp {g}: sigma {sa} vs {sb}"
        assert np.array_equal(ca, cb), f"{where} gp {g}: ctan"
        assert fa == fb, f"{where} gp {g}: flags {fa} vs {fb}"
        for w in range(4):
            assert (xa[w] is None) == (xb[w] is None), f"{where} gp {g}: state {w} presence"
            if xa[w] is not None:
                assert np.array_equal(xa[w], xb[w]), f"{where} gp {g}: state {w}"


KW = dict(subiterations=1, nsubiterations=3, nr_max_its=3, nr_rel_tol=1e-6, cg_rel_tol=1e-10, cg_max_its=3000)


@pytest.mark.parametrize("lin_stress,its_with_A0", [(0, 0), (1, 0), (0, 1), (0, 2)])
def test_gp_batch_equals_one_gp_path(mp, monkeypatch, lin_stress, its_with_A0):
    """its_with_A0 > 0: Newton on the initial matrix (use_A0) in both paths."""
    steps = 3
    a0 = dict(use_A0=1, its_with_A0=its_with_A0) if its_with_A0 else {}
    seq, grp, one = (make(mp, lin_stress=lin_stress, **KW, **a0) for _ in range(3))
    rp = ramps(seq.ngp, steps)
    nl = sub = 0
    for t in range(steps):
        a = step(seq, rp, t, monkeypatch, 0)
        b = step(grp, rp, t, monkeypatch, 2)
        c = step(one, rp, t, monkeypatch, -1)  # no cap: every FE GP in one batch
        print(f"step {t}: flags " + " ".join(str(r[2]) for r in a), flush=True)
        same(a, b, f"step {t} (2 slots)")
        same(a, c, f"step {t} (5 slots)")
        nl += sum(r[2][0] for r in a)
        sub += sum(r[2][3] for r in a)
    assert nl > 0 and sub > 0, "the case must exercise history allocation and subiterations"


def test_gp_batch_after_restart(mp, monkeypatch, tmp_path):
    """History read from a restart file feeds the batched path as it feeds the one-GP path."""
    monkeypatch.chdir(tmp_path)
    steps = 3
    rp = ramps(len(couplings(mp)), steps, seed=9)
    a = make(mp, lin_stress=0, **KW)
    for t in range(steps - 1):
        step(a, rp, t, monkeypatch, 2)
    a.write_restart(5)
    b, c = make(mp, lin_stress=0, **KW), make(mp, lin_stress=0, **KW)
    b.read_restart(5)
    c.read_restart(5)
    nl = 0
    for g in range(a.ngp):
        assert b.is_non_linear(g) == a.is_non_linear(g)
        if a.is_non_linear(g):
            nl += 1
            assert np.array_equal(b.state(g, 0), a.state(g, 0)) and np.array_equal(b.state(g, 2), a.state(g, 2))
    assert nl > 0
    rb = step(b, rp, steps - 1, monkeypatch, 2)
    rc = step(c, rp, steps - 1, monkeypatch, 0)
    same(rb, rc, "after restart")


def test_gp_batch_matches_oracle(mp, monkeypatch, oracle):
    steps = 2
    kw = dict(lin_stress=0, **KW)
    m = make(mp, **kw)
    cp = couplings(mp)
    p, k = ob.make_params(6, mp.MIC_SPHERE, (0.3, 0, 0, 0), [ob.material(*s) for s in J2], ngp=len(cp), coupling=cp,
                          **kw)
    o = oracle.micro(p, k)
    rp = ramps(len(cp), steps)
    for t in range(steps):
        r = step(m, rp, t, monkeypatch, 2)
        for g in range(len(cp)):
            o.set_strain(g, rp[g][t])
        o.homogenize()
        for g in range(len(cp)):
            s0, c0 = o.get_stress(g), o.get_ctan(g)
            e = np.abs(r[g][0] - s0).max() / max(np.abs(s0).max(), 1e-300)
            ec = np.abs(r[g][1] - c0).max() / max(np.abs(c0).max(), 1e-300)
            print(f"step {t} gp {g}: sigma rel {e:.2e} ctan rel {ec:.2e}", flush=True)
            assert e < 1e-6 and ec < 1e-6
            assert (r[g][2][0], r[g][2][2], r[g][2][3]) == (o.is_non_linear(g), int(o.has_converged(g)),
                                                            int(o.has_subiterated(g)))
        o.update_vars()
