"""Batched density / Jacobi and the request loop on the GPU: every molecule's
result in a batch equals its result alone; graph replays equal eager runs."""
import time

import numpy
import pytest
from conftest import BASIS

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def items():
    import qmtools
    from qmtools import synth
    basis = qmtools.BasisSet(BASIS)
    out = []
    for natoms, seed, n, step in [(5, 1, 24, 0.45), (9, 2, 32, 0.4), (14, 3, 16, 0.8), (7, 4, 40, 0.3)]:
        mol = synth.synthetic_molecule(natoms, seed, basis)
        out.append((mol, synth.centred_grid(mol, n, step_bohr=step, pad_ang=2.0)))
    return out


@pytest.mark.parametrize("precision,knob,variant", [("fp64", "f64_variant", 2), ("fp64", "f64_variant", 0),
                                                    ("fp64", "f64_variant", 3), ("fp64", "f64_variant", 4),
                                                    ("mixed", "split_variant", 0), ("mixed", "split_variant", 1),
                                                    ("mixed", "split_variant", 2), ("mixed", "split_variant", 3),
                                                    ("mixed", "split_variant", 4), ("mixed", "split_variant", 128)])
def test_density_batch_equals_single(items, precision, knob, variant):
    import qmtools
    from qmtools import autotune, batch
    mols = [m for m, _ in items]
    grids = [g for _, g in items]
    _say("density batch", precision, knob, variant, [g.shape for g in grids])
    old = autotune.get(knob)
    autotune.set(knob, variant)
    try:
        qs = batch.density_batch(mols, grids, precision)
        for (m, g), q in zip(items, qs):
            single = qmtools.QMTools.Compute_density(g, m, precision=precision)
            numpy.testing.assert_array_equal(q.cpu().numpy(), single.qube.ravel())
    finally:
        autotune.set(knob, old)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fft_batch_equals_single(items, precision):
    """Batched FFT Poisson equals the drop-in call molecule by molecule (FP64 bitwise; FP32 <= 1e-6 vs FP64)."""
    import qmtools
    from qmtools import batch
    mols = [m for m, _ in items]
    grids = [g for _, g in items]
    _say("fft batch", precision, [g.shape for g in grids])
    qs, vs = batch.run(mols, grids, precision, poisson="fft")
    for (m, g), q, v in zip(items, qs, vs):
        qg = qmtools.Grid.emptyAs(g)
        qg.d_qube = q.double().clone()
        want = qmtools.QMTools.Compute_hartree_fft(qg, m, mode="freespace", precision=precision)
        got = v.double().cpu().numpy()
        got = got.reshape(g.space_shape())           # both are V[i, j, k] in C order
        if precision == "fp64":
            numpy.testing.assert_array_equal(got, want)
        else:
            ref = qmtools.QMTools.Compute_hartree_fft(qg, m, mode="freespace", precision="fp64")
            assert numpy.abs(got - ref).max() <= 1e-6 * numpy.abs(ref).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_jacobi_batch_equals_single(items, precision):
    import qmtools
    from qmtools import batch
    mols = [m for m, _ in items]
    grids = [g for _, g in items]
    sweeps = 4 * (100 if precision == "fp64" else 50) + 7      # several graph replays + an odd remainder
    _say("jacobi batch", precision, sweeps)
    qs = batch.density_batch(mols, grids, precision)
    vs_graph = batch.hartree_jacobi_batch(mols, grids, qs, sweeps, precision, "multipole", use_graph=True)
    vg1 = [v.cpu().numpy() for v in vs_graph]
    # eager work and a synchronisation between the two graph runs
    vs_eager = batch.hartree_jacobi_batch(mols, grids, qs, sweeps, precision, "multipole", use_graph=False)
    ve = [v.cpu().numpy() for v in vs_eager]
    vs_graph2 = batch.hartree_jacobi_batch(mols, grids, qs, sweeps, precision, "multipole", use_graph=True)
    vg2 = [v.cpu().numpy() for v in vs_graph2]
    for a, b, c in zip(vg1, ve, vg2):
        numpy.testing.assert_array_equal(a, b)
        numpy.testing.assert_array_equal(c, b)
    if precision == "fp64":
        for (m, g), q, v in zip(items, qs, ve):
            qg = qmtools.Grid.emptyAs(g)
            qg.d_qube = q.clone()
            single = qmtools.QMTools.Compute_hartree_iter(qg, m, tolerance=None, max_sweeps=sweeps,
                                                          boundary="multipole")
            numpy.testing.assert_allclose(v, single.qube.ravel(), rtol=0, atol=1e-13 * numpy.abs(v).max())


def test_jacobi_batch_mixed_vs_oracle(items):
    from oracle import jacobi as oj
    from oracle import seed as osd
    from qmtools import batch
    mols = [m for m, _ in items]
    grids = [g for _, g in items]
    _say("jacobi batch mixed vs oracle 2000")
    qs = batch.density_batch(mols, grids, "mixed")
    vs = batch.hartree_jacobi_batch(mols, grids, qs, 2000, "mixed", "multipole")
    for (m, g), q, v in zip(items, qs, vs):
        nx, ny, nz, h, o = g.geometry()
        rho = osd.seed_grid(m.types, m.coords, nx, ny, nz, h, o) - q.double().cpu().numpy()
        want = oj.jacobi(rho, nx, ny, nz, h, 2000, bvals=oj.multipole_boundary(rho, nx, ny, nz, h, o))
        err = numpy.abs(v.cpu().numpy() - want).max() / numpy.abs(want).max()
        assert err <= 1e-6, err


def test_server_results_same_alone_and_in_mix(items):
    from qmtools import batch
    from qmtools.serve import Server
    _say("server mix vs alone")
    srv = Server(nlanes=2, precision="mixed", max_batch=2, sweeps=300).start()
    reqs = [srv.submit(m, g) for m, g in items for _ in range(2)]
    for r in reqs:
        r.wait(120)
    srv.stop()
    assert srv.max_concurrent >= 1
    for r in reqs:
        qs, vs = batch.run([r.mol], [r.grid], "mixed", 300)
        numpy.testing.assert_array_equal(r.density, qs[0].cpu().numpy())
        numpy.testing.assert_array_equal(r.potential, vs[0].cpu().numpy())
    st = srv.stats()
    assert st["requests"] == len(reqs) and st["failed"] == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_server_rounds_late_joiners_equal_alone(items, precision):
    """Jacobi lanes in rounds of 100 sweeps: three requests submitted after the first was admitted
    join its lane at a round boundary, and every request's density and potential equal those of
    the molecule run alone to the same sweep count, bit for bit."""
    from qmtools import batch
    from qmtools.serve import Server
    _say("server rounds", precision)
    srv = Server(nlanes=1, precision=precision, max_batch=4, sweeps=1000, round_sweeps=100).start()
    first = srv.submit(*items[3])
    while first.t_admit is None:
        time.sleep(1e-3)
    late = [srv.submit(m, g) for m, g in items[:3]]
    reqs = [first] + late
    for r in reqs:
        r.wait(120)
    srv.stop()
    assert srv.joined_running >= 1
    assert min(r.t_admit for r in late) < first.t_done
    for r in reqs:
        qs, vs = batch.run([r.mol], [r.grid], precision, 1000)
        numpy.testing.assert_array_equal(r.density, qs[0].cpu().numpy())
        numpy.testing.assert_array_equal(r.potential, vs[0].cpu().numpy())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_jacobi_rounds_alternating_sets_equal_one_call(items, precision):
    """The batched Jacobi keeps one block graph per context and reuses it while the item table is
    unchanged: sweeping set A, then set B, then A again (each in calls of whole blocks) must give
    every molecule the bits of one call with all its sweeps."""
    from qmtools import batch
    _say("jacobi alternating sets", precision)
    lane = batch.Lane()
    blk = batch.jacobi_block(precision)
    mols = [m for m, _ in items]
    grids = [g for _, g in items]
    qs = batch.density_batch(mols, grids, precision, lane)
    sets = {"A": [0, 1], "B": [2, 3, 0]}
    states = {k: batch.jacobi_prepare([mols[i] for i in v], [grids[i] for i in v], [qs[i] for i in v],
                                      precision, "multipole", lane) for k, v in sets.items()}
    done = {"A": 0, "B": 0}
    for k in ("A", "B", "A", "B", "A"):
        batch.jacobi_sweeps([grids[i] for i in sets[k]], states[k], 2 * blk, precision, lane)
        done[k] += 2 * blk
    lane.stream.synchronize()
    for k, v in sets.items():
        want = batch.hartree_jacobi_batch([mols[i] for i in v], [grids[i] for i in v], [qs[i] for i in v],
                                          done[k], precision, "multipole", lane)
        lane.stream.synchronize()
        for st, w in zip(states[k], want):
            numpy.testing.assert_array_equal(st[1].cpu().numpy(), w.cpu().numpy())
    lane.close()


def test_jacobi_fp32_paired_columns_bitwise(items):
    """FP32 two-sweep passes with two columns per lane equal the one-column kernel bit for bit: single
    grids (one strip, and 136 x 24 x 16 with two strips and a part-filled last one) and a batch, over
    several graph replays and an odd remainder of sweeps."""
    import qmtools
    from qmtools import autotune, batch, synth
    from qmtools.grid import Grid, make_float3
    m0, g0 = items[0]
    wide = Grid((1, 136, 24, 16), 1.0)
    wide.step = numpy.float64(0.35)
    c = m0.coords.mean(axis=0)
    wide.origin = make_float3(*(c - 0.5 * 0.35 * numpy.array([136, 24, 16])))
    cases = [(m0, g0), (m0, wide)]
    sweeps = 3 * 50 + 7
    old = autotune.get("jacobi_pairs")
    out = {}
    try:
        for pairs in (1, 0):
            autotune.set("jacobi_pairs", pairs)
            res = []
            for m, g in cases:
                _say("jacobi fp32 pairs", pairs, g.shape, sweeps)
                q = qmtools.QMTools.Compute_density(g, m, precision="mixed", copyBack=False)
                res.append(qmtools.QMTools.Compute_hartree_iter(q, m, tolerance=None, max_sweeps=sweeps,
                                                                boundary="multipole", precision="mixed").qube.ravel())
            mols = [m for m, _ in items] + [m0]
            grids = [g for _, g in items] + [wide]
            _say("jacobi fp32 batch pairs", pairs, [g.shape for g in grids])
            qs = batch.density_batch(mols, grids, "mixed")
            res += [v.cpu().numpy() for v in batch.hartree_jacobi_batch(mols, grids, qs, sweeps, "mixed",
                                                                          "multipole")]
            out[pairs] = res
    finally:
        autotune.set("jacobi_pairs", old)
    assert len(out[1]) == len(out[0]) == 2 + len(items) + 1
    for a, b in zip(out[1], out[0]):
        assert numpy.abs(b).max() > 0
        numpy.testing.assert_array_equal(a, b)
