"""One evaluation (update_density -> update_vhxc -> energy with H*psi,
forces and stress) captured in a HIP graph: every replay, with eager work in
between, gives exactly the eager results; several systems in one graph; calls
that synchronize are refused while capturing."""
import numpy as np
import pytest

from qbg import configs as cf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import qbg
    c = qbg.Context(0)
    yield c
    c.close()


def _same(a, b):
    for k in ("etotal", "ekin", "eps", "enl", "ehart", "exc"):
        assert a[k] == b[k], (k, a[k], b[k])
    assert np.array_equal(a["fion"], b["fion"])
    assert np.array_equal(a["sigma"], b["sigma"])


@pytest.mark.parametrize("idx,prec", [(11, "fp64"), (0, "fp64"), (0, "bf16x3")])
def test_graph_replay_matches_eager(ctx, idx, prec):
    import qbg
    s = cf.config(idx)
    print(f"[case] config {idx} {prec}", flush=True)
    ef = qbg.EnergyFunctional(ctx, s)
    b = ef.basis(0)
    c = cf.gram_numpy(cf.random_coefficients(900 + idx, b["g2"], s.nst))
    c2 = cf.gram_numpy(cf.random_coefficients(901 + idx, b["g2"], s.nst))
    dc, dc2 = ctx.to_device(c), ctx.to_device(c2)
    hp = ctx.array(c.shape, np.complex128)
    ref = ef.evaluate(dc, hp, forces=True, stress=True, prec=prec)
    href = hp.download()
    # first (uncaptured) async call allocates what the graph will use
    ef.evaluate_async(dc, hp, forces=True, stress=True, prec=prec)
    _same(ef.results(), ref)
    with qbg.Graph(ctx) as g:
        ef.evaluate_async(dc, hp, forces=True, stress=True, prec=prec)
    for rep in range(3):
        hp.zero()
        g.launch()
        got = ef.results()
        _same(got, ref)
        assert np.array_equal(hp.download(), href)
        # eager work on other inputs between replays
        other = ef.evaluate(dc2, None, forces=True, stress=False, prec=prec)
        assert other["etotal"] != ref["etotal"]
    g.release()


def test_graph_several_systems(ctx):
    import qbg
    efs, dcs, refs = [], [], []
    for idx in (10, 11, 12):
        s = cf.config(idx)
        ef = qbg.EnergyFunctional(ctx, s)
        c = cf.gram_numpy(cf.random_coefficients(700 + idx, ef.basis(0)["g2"], s.nst))
        dc = ctx.to_device(c)
        refs.append(ef.evaluate(dc, None, forces=True))
        ef.evaluate_async(dc, None, forces=True)
        ef.results()
        efs.append(ef)
        dcs.append(dc)
    with qbg.Graph(ctx) as g:
        for ef, dc in zip(efs, dcs):
            ef.evaluate_async(dc, None, forces=True)
    for rep in range(2):
        g.launch()
        for ef, ref in zip(efs, refs):
            got = ef.results()
            assert got["etotal"] == ref["etotal"]
            assert np.array_equal(got["fion"], ref["fion"])
    g.release()


def test_capture_refuses_sync(ctx):
    import qbg
    s = cf.config(10)
    ef = qbg.EnergyFunctional(ctx, s)
    dc = ctx.to_device(cf.gram_numpy(cf.random_coefficients(5, ef.basis(0)["g2"], s.nst)))
    ef.evaluate(dc, None, forces=False)
    with pytest.raises(qbg.QbgError) as e:
        with qbg.Graph(ctx):
            ef.evaluate_async(dc, None, forces=False)
            ef.results()              # waits on the stream: refused while capturing
    assert e.value.code == -5


def test_parallel_branches_equal_eager():
    """Several systems, one context (stream) each, captured as parallel
    branches of one graph (qbg_stream_join): every replay gives each system's
    eager results bit for bit, also after other work ran in between."""
    import qbg
    main = qbg.Context(0)
    parts = []
    for idx in (10, 11, 12, 0):
        s = cf.config(idx)
        ctx = qbg.Context(0)
        ef = qbg.EnergyFunctional(ctx, s)
        c = cf.gram_numpy(cf.random_coefficients(600 + idx, ef.basis(0)["g2"], s.nst))
        dc = ctx.to_device(c)
        hp = ctx.array(c.shape, np.complex128)
        prec = "bf16x2" if idx == 0 else "fp64"
        ref = ef.evaluate(dc, hp, forces=True, prec=prec)
        parts.append((ctx, ef, dc, hp, prec, ref, hp.download()))
    print("[case] parallel graph branches", flush=True)
    with qbg.Graph(main) as g:
        for ctx, ef, dc, hp, prec, _, _ in parts:
            qbg.stream_join(main, ctx)
            ef.evaluate_async(dc, hp, forces=True, prec=prec)
        for ctx, *_ in parts:
            qbg.stream_join(ctx, main)
    for rep in range(3):
        for ctx, _, _, hp, *_ in parts:
            hp.zero()
            ctx.sync()             # the graph runs on main's stream
        g.launch()
        main.sync()
        for ctx, ef, dc, hp, prec, ref, href in parts:
            got = ef.results()
            assert got["etotal"] == ref["etotal"], rep
            assert np.array_equal(got["fion"], ref["fion"])
            assert np.array_equal(hp.download(), href)
        # eager work on one branch's context between replays
        parts[1][1].evaluate(parts[1][2], parts[1][3], forces=True)
    g.release()
    for ctx, ef, *_ in parts:
        ef.release()
        ctx.close()
    main.close()
