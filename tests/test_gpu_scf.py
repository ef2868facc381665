"""The device SCF-iteration updates against the oracle: DoubleMatrix::polar
in both of its branches, single updates of the SD, PSD and PSDA steppers
(PSDA twice, so the second extrapolates), the charge mixer over several
iterations (Anderson and simple), and whole SCF loops run eagerly and from
replayed graphs.  FP64 throughout; the tolerances allow for the different
summation order of the MFMA GEMMs and block reductions."""
import numpy as np
import pytest

from qbg import configs as cf

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("n,scale", [(9, 0.05), (9, 2.0), (200, 0.02), (130, 1.0)])
def test_polar(ctx, oracle_mod, n, scale):
    import qbg
    print(f"[case] polar n={n} scale={scale}", flush=True)
    rng = np.random.default_rng(n)
    x = np.eye(n) + scale * rng.standard_normal((n, n)) / np.sqrt(max(1, n // 9))
    u, it = qbg.polar(ctx, x)
    uo, ito = oracle_mod.polar(x)
    W, _, Vt = np.linalg.svd(x)
    assert it == ito
    assert np.abs(u - uo).max() < 1e-10
    assert np.abs(u - W @ Vt).max() < 1e-10


def _hpsi(oracle_mod, o, c):
    r = o.evaluate(c, hpsi=True, forces=False)
    return r["hpsi"]


@pytest.mark.parametrize("idx,kind,param,nst", [(11, "sd", 0.3, None), (11, "psd", 0.0, None),
                                                (12, "psd", 2.0, None), (11, "psda", 0.0, None),
                                                (0, "psda", 0.0, None), (1, "psda", 0.0, 160),
                                                (1, "psd", 0.0, 150)])
def test_stepper_update(ctx, oracle_mod, idx, kind, param, nst):
    """nst above 128 runs the device-wide blocked Cholesky inside gram and
    ortho_align."""
    import dataclasses
    import qbg
    print(f"[case] stepper {kind} config {idx} nst {nst}", flush=True)
    s = cf.config(idx)
    if nst is not None:
        s = dataclasses.replace(s, nst=nst)
    ef = qbg.EnergyFunctional(ctx, s)
    o = oracle_mod.Oracle(s)
    b = ef.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(40 + idx, b["g2"], s.nst))
    st = qbg.WfStepper(ef, kind, param)
    so = oracle_mod.WfStepper(kind, param, b["g2"])
    dc = ctx.to_device(c)
    dh = ctx.array(c.shape, np.complex128)
    co = c.copy()
    for step in range(2 if kind == "psda" else 1):
        h = _hpsi(oracle_mod, o, co)
        eig = oracle_mod.sd_dot(co, h)
        dh.upload(h)
        st.update(dc, dh)
        w = st.results()
        co, ho = so.update(co, h)
        assert abs(w["eigenvalue_sum"] - eig) <= TOL * abs(eig)
        assert w["status"] == 0
        cg, hg = dc.download(), dh.download()
        assert rel(hg, ho) < 1e-9, (step, rel(hg, ho))
        assert rel(cg, co) < TOL, (step, rel(cg, co))
        # continue both from the oracle's bands so the comparison stays per step
        dc.upload(co)
    if kind == "psda":
        assert w["polar_iterations"] >= 1
    st.release()
    ef.release()


@pytest.mark.parametrize("ndim", [3, 0])
def test_mixer_sequence(ctx, oracle_mod, ndim):
    import qbg
    print(f"[case] mixer ndim={ndim}", flush=True)
    s = cf.config(11)
    ef = qbg.EnergyFunctional(ctx, s)
    o = oracle_mod.Oracle(s)
    b = ef.basis(0)
    mx = qbg.Mixer(ef, 0.4, 6.0, ndim, 3.0)
    mo = o.mixer(0.4, 6.0, ndim, 3.0)
    for it in range(6):
        c = cf.gram_numpy(cf.random_coefficients(70 + it, b["g2"], s.nst))
        dc = ctx.to_device(c)
        ef.update_density(dc)
        mx.update()
        o.update_density(c)
        mo.update()
        assert rel(ef.get(3), o.get(3)) < 1e-12, it
        assert rel(ef.get(0), o.get(0)) < 1e-12, it
        dc.release()
    mx.release()
    ef.release()


def _device_scf(ctx, s, c, kind, param, nitscf, nite, graphs, ndim=3):
    import qbg
    from qbg.scf import ScfLoop
    ef = qbg.EnergyFunctional(ctx, s)
    dc = ctx.to_device(c)
    dh = ctx.array(c.shape, np.complex128)
    loop = ScfLoop(ef, dc, dh, wf_dyn=kind, param=param, ndim=ndim, graphs=graphs)
    hist = loop.run(nitscf, nite)
    cf_ = dc.download()
    loop.release()
    ef.release()
    return hist, cf_


@pytest.mark.parametrize("kind,param", [("psda", 0.0), ("psd", 0.0), ("sd", 0.25)])
def test_scf_loop_matches_oracle(ctx, oracle_mod, kind, param):
    from test_oracle_scf import scf
    print(f"[case] scf loop {kind}", flush=True)
    s = cf.config(10)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(21, b["g2"], s.nst))
    hist, cdev = _device_scf(ctx, s, c, kind, param, 6, 1, graphs=False)
    co, es = scf(oracle_mod, s, o, c, kind, param, 6, 1)
    ed = [h["etotal"] for h in hist]
    assert np.allclose(ed, es, rtol=1e-10, atol=0), (ed, es)
    assert rel(cdev, co) < 1e-8


def test_scf_graph_replay_equals_eager(ctx, oracle_mod):
    print("[case] scf graphs", flush=True)
    s = cf.config(0)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(22, b["g2"], s.nst))
    h1, c1 = _device_scf(ctx, s, c, "psda", 0.0, 5, 2, graphs=False)
    h2, c2 = _device_scf(ctx, s, c, "psda", 0.0, 5, 2, graphs=True)
    for a, bb in zip(h1, h2):
        assert a["etotal"] == bb["etotal"]
        assert a["eigenvalue_sum"] == bb["eigenvalue_sum"]
    assert np.array_equal(c1, c2)


def test_scf_psda_converges(ctx, oracle_mod):
    """The device loop reaches the minimum the oracle's loop finds
    (-9.0401854 Ha for this system)."""
    print("[case] scf converge", flush=True)
    s = cf.config(10)
    o = oracle_mod.Oracle(s)
    b = o.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(21, b["g2"], s.nst))
    hist, cdev = _device_scf(ctx, s, c, "psda", 0.0, 40, 4, graphs=True)
    e = [h["etotal"] for h in hist]
    assert abs(e[-1] - e[-2]) < 1e-6
    o.update_density(cdev)
    o.update_vhxc(False)
    eo = o.energy(cdev, hpsi=False, forces=False)["etotal"]
    assert abs(eo - (-9.0401854)) < 2e-6, eo
