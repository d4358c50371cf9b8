"""The oracle split over worker processes (tests/oracle_pool.py) gives the
oracle's own results: tiles and layer ranges of the forward, layer ranges of
the adjoint, and the band-mask line filter sent to each task."""

import numpy as np

from oracle import adjoint_oracle as A
from oracle import lbl_oracle as O
from oracle_pool import adjoint_vjp, executor, lines_near, optical_depth_tiles
from pylblrtm_amd import synth
from test_adjoint_oracle import translucent_case


def test_lines_near_keeps_every_line_that_reaches_the_points():
    rng = np.random.default_rng(3)
    lines = synth.random_lines(rng, 4000, 600.0, 800.0, 2)
    lines["shift"] = rng.normal(0.0, 0.05, 4000)
    near = lines_near(lines, 700.0, 702.0, 25.0, 1.0)
    for P in (0.1, 0.5, 1.0):
        c = lines["nu0"] + lines["shift"] * P
        reach = (c >= 700.0 - 25.0) & (c <= 702.0 + 25.0)
        assert set(np.asarray(lines["nu0"])[reach]) <= set(near["nu0"])
    assert 0 < near["nu0"].size < 4000


def test_split_forward_equals_the_oracle():
    sd, prof = synth.five_gas(lines_per_gas=300, n_levels=7, nu_lo=900.0, nu_hi=960.0)
    lay = prof.build_layers()
    args = (sd.lines, sd.tips, sd.molar_masses, lay.pressure_atm, lay.temperature_k, lay.vmr,
            lay.column_cm2)
    grid = (920.0, 930.0, 0.01)
    nu = O.grid(*grid)
    tiles = [0, 2, 5]
    ref = O.optical_depth(nu, *args, 25.0, 2.0, tiles=tiles)
    with executor(2) as pool:
        got = optical_depth_tiles(pool, grid, tiles, *args, 25.0, 2.0, chunks_per_tile=3)
    ts = O.tile_samples(nu, 2.0)
    for tl in tiles:
        a, b = tl * ts, min(tl * ts + ts, nu.size)
        np.testing.assert_allclose(got[tl], ref[:, a:b], rtol=1e-13, atol=0)
    assert got[5].shape[1] == 1            # the last, one-point tile


def test_split_adjoint_equals_the_oracle():
    nu, lines, tips, mm, p, t, vmr, air = translucent_case()
    mus = (1.0, 0.5)
    G = np.random.default_rng(2).normal(size=(2, nu.size))
    rt, rv = A.vjp(nu, lines, tips, mm, p, t, vmr, air, 25.0, 0.5, mus, G)
    with executor(2) as pool:
        gt, gv, Tm = adjoint_vjp(pool, (700.0, 702.0, 0.004), lines, tips, mm, p, t, vmr, air,
                                 25.0, 0.5, mus, G)
    np.testing.assert_allclose(Tm, A.forward(nu, lines, tips, mm, p, t, vmr, air, 25.0, 0.5, mus),
                               rtol=1e-13)
    np.testing.assert_allclose(gt, rt, rtol=1e-10)
    for g in rv:
        np.testing.assert_allclose(gv[g], rv[g], rtol=1e-10)
