"""GPU parity of the HIP step against the oracle (the CPU restatement that is
itself bit-exact against the reference's CPU sources, see test_oracle_ref.py).

The exact variant (-ffp-contract=off kernels) must agree with the oracle value
for value (np.array_equal; -0.0 == 0.0) on every field, ghosts included, and on
dt/time/nstep. The FMA variant must agree to L1-relative <= 1e-12 after 100
steps on every conserved and face-B field, the floating-point parity
tolerance of the path.
"""
import os

import numpy as np
import pytest

from helpers import INPUTS, interior, l1_rel, load_pkg, oracle_setup

pytestmark = pytest.mark.gpu

FIELDS = ("gas", "B1i", "B2i", "B3i")


def say(msg):
    print(msg, flush=True)


def compare_exact(gpu, ora, tag):
    a = gpu.download(0)
    b = ora.get()
    for name, x, y in zip(FIELDS, a, b):
        if not np.array_equal(x, y):
            diff = np.abs(x - y)
            idx = np.unravel_index(np.argmax(diff), diff.shape)
            raise AssertionError(f"{tag}: {name} differs, max |diff| {diff.max():.3e} at {idx}")
    dt_g, t_g, n_g = gpu.clock(0)
    dt_o, t_o, n_o = ora.clock()
    assert (dt_g, t_g, n_g) == (dt_o, t_o, n_o), f"{tag}: clock {dt_g, t_g, n_g} vs {dt_o, t_o, n_o}"


CASES = [
    ("field_loop", [], 40),
    ("blast", ["grid/Nx1=64", "grid/Nx2=96"], 40),
    ("cpaw2d", ["grid/Nx1=32", "grid/Nx2=32"], 30),
]


# kernel forms: (x2 predictor form, fusion depth); results must not depend on them
FORMS = [(0, 0), (1, 0), (1, 1), (1, 2)]


@pytest.mark.parametrize("forms", FORMS, ids=lambda f: f"x2form{f[0]}_fusion{f[1]}")
@pytest.mark.parametrize("flux", ["roe", "hlle"])
@pytest.mark.parametrize("problem,overrides,nsteps", CASES)
def test_exact_steps_match_oracle(problem, overrides, nsteps, flux, forms):
    pkg = load_pkg()
    say(f"[gpu] exact {problem} {flux} {overrides} steps={nsteps} forms {forms}")
    gpu, par, tlim, nlim = pkg.setup_problem(problem, overrides, flux=flux, variant="exact")
    gpu.set_x2_kind(forms[0])
    gpu.set_fusion(forms[1])
    ora, tlim_o, _ = oracle_setup(problem, overrides, flux=flux)
    compare_exact(gpu, ora, "initial state")
    gpu.set_limits(tlim, nsteps)
    done = 0
    for chunk in (1, 4, nsteps - 5):
        for _ in range(chunk):
            gpu.step()
            ora.step(tlim)
        done += chunk
        compare_exact(gpu, ora, f"after {done} steps")
    gpu.close()


@pytest.mark.parametrize("bcs,fusion", [((2, 2, 2, 2), 1), ((1, 1, 1, 1), 2), ((1, 2, 4, 4), 2),
                                        ((4, 4, 2, 1), 1)])
def test_exact_boundaries_match_oracle(bcs, fusion):
    pkg = load_pkg()
    ov = ["grid/Nx1=48", "grid/Nx2=40", f"grid/ibc_x1={bcs[0]}", f"grid/obc_x1={bcs[1]}",
          f"grid/ibc_x2={bcs[2]}", f"grid/obc_x2={bcs[3]}", "problem/radius=0.25"]
    say(f"[gpu] exact blast boundaries {bcs} fusion {fusion}")
    gpu, par, tlim, _ = pkg.setup_problem("blast", ov, variant="exact")
    gpu.set_fusion(fusion)
    ora, _, _ = oracle_setup("blast", ov)
    compare_exact(gpu, ora, "initial ghosts")
    gpu.set_limits(tlim, 25)
    for s in range(25):
        gpu.step()
        ora.step(tlim)
    compare_exact(gpu, ora, f"boundaries {bcs} after 25 steps")
    gpu.close()


@pytest.mark.parametrize("rows", [1, 2, 4])
@pytest.mark.parametrize("nx1,nx2", [(61, 61), (64, 62), (125, 29)])
def test_fused_corner_tile_edges(nx1, nx2, rows):
    """Fusion depth 2: the x1 corrector's tiles compute the Step 4 corners and
    store those they own. Grids where the corner range's last row or column
    (je+2, ie+2) falls just outside the last tile's own rows or columns
    ((nx + 3) a multiple of the tile size) or inside them; both fluxes."""
    pkg = load_pkg()
    ov = [f"grid/Nx1={nx1}", f"grid/Nx2={nx2}", "problem/radius=0.2"]
    for flux in ("roe", "hlle"):
        say(f"[gpu] fusion 2 blast {nx1}x{nx2} corrector rows {rows} {flux}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, flux=flux, variant="exact")
        gpu.set_fusion(2)
        gpu.set_corr_rows(rows)
        ora, _, _ = oracle_setup("blast", ov, flux=flux)
        gpu.set_limits(tlim, 12)
        for s in range(12):
            gpu.step()
            ora.step(tlim)
        compare_exact(gpu, ora, f"fusion 2 {nx1}x{nx2} rows {rows} {flux} after 12 steps")
        gpu.close()


@pytest.mark.parametrize("seg", [1, 3, 5, 7, 24])
def test_x2_marching_depths(seg):
    """x2 predictor segments of 1..24 faces per lane (athmi_set_shape), both
    forms: 73 rows of cells, so 76 rows of x2 faces (js-1 .. je+2) and the
    last segment is partial for every depth but 1; a persistent grid beside
    one tile per workgroup."""
    pkg = load_pkg()
    ov = ["grid/Nx1=150", "grid/Nx2=73", "problem/radius=0.2"]
    for kind, persist, flux in ((1, 0, "roe"), (0, 0, "hlle"), (1, 4, "hlle"), (0, 2, "roe")):
        say(f"[gpu] x2 depth {seg} form {kind} persist {persist} blast 150x73 {flux}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, flux=flux, variant="exact")
        gpu.set_x2_kind(kind)
        gpu.set_shape(seg, persist)
        assert gpu.launch_shape()[:2] == (seg, persist)
        ora, _, _ = oracle_setup("blast", ov, flux=flux)
        gpu.set_limits(tlim, 10)
        for s in range(10):
            gpu.step()
            ora.step(tlim)
        compare_exact(gpu, ora, f"x2 depth {seg} form {kind} persist {persist} {flux} after 10 steps")
        gpu.close()


@pytest.mark.parametrize("strip", [2, 3, 4, 8])
def test_strip_tile_order(strip):
    """Corrector and final tiles walked in strips of `strip` tile columns
    (athmi_set_strip): five tile columns, so the strips are uneven (2+2+1,
    3+2, 4+1) or wider than the grid (8: row order); every corrector tile
    height and fusion depth 1 and 2."""
    pkg = load_pkg()
    ov = ["grid/Nx1=300", "grid/Nx2=72", "problem/radius=0.2"]
    for rows, fusion in ((1, 1), (2, 2), (4, 1), (4, 2)):
        say(f"[gpu] strip {strip} blast 300x72 corrector rows {rows} fusion {fusion}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, variant="exact")
        gpu.set_strip(strip)
        gpu.set_corr_rows(rows)
        gpu.set_fusion(fusion)
        assert gpu.strip() == strip
        ora, _, _ = oracle_setup("blast", ov)
        gpu.set_limits(tlim, 10)
        for s in range(10):
            gpu.step()
            ora.step(tlim)
        compare_exact(gpu, ora, f"strip {strip} rows {rows} fusion {fusion} after 10 steps")
        gpu.close()


@pytest.mark.parametrize("rows", [4, 8, 16])
def test_final_tile_rows(rows):
    """Final-update tiles of 64 x 4, 8 or 16 cells (athmi_set_final_rows):
    73 rows of faces, so the last tile row is partial at every height; row by
    row and in uneven strips, fusion depth 0 and 1."""
    pkg = load_pkg()
    ov = ["grid/Nx1=300", "grid/Nx2=72", "problem/radius=0.2"]
    for strip, fusion in ((0, 1), (3, 1), (0, 0)):
        say(f"[gpu] final rows {rows} blast 300x72 strip {strip} fusion {fusion}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, variant="exact")
        gpu.set_final_rows(rows)
        gpu.set_strip(strip)
        gpu.set_fusion(fusion)
        assert gpu.final_rows() == rows
        ora, _, _ = oracle_setup("blast", ov)
        gpu.set_limits(tlim, 10)
        for s in range(10):
            gpu.step()
            ora.step(tlim)
        compare_exact(gpu, ora, f"final rows {rows} strip {strip} fusion {fusion} after 10 steps")
        gpu.close()


@pytest.mark.parametrize("bcs", [(4, 4, 4, 4), (2, 2, 2, 2), (1, 1, 1, 1), (1, 2, 4, 4)])
@pytest.mark.parametrize("nx1,nx2", [(2, 2), (3, 5), (4, 4), (2, 64), (64, 2), (4, 40), (40, 4),
                                     (37, 53)])
def test_exact_tiny_and_thin_grids_match_oracle(nx1, nx2, bcs):
    """Grids thinner than the four ghost layers (the reference's ghost fill
    then copies ghost cells it wrote earlier in the same pass), thin strips
    and odd sizes, every boundary kind, both fluxes, stepped as the main loop
    does until tlim or 12 steps."""
    pkg = load_pkg()
    ov = [f"grid/Nx1={nx1}", f"grid/Nx2={nx2}", "problem/radius=0.2",
          f"grid/ibc_x1={bcs[0]}", f"grid/obc_x1={bcs[1]}", f"grid/ibc_x2={bcs[2]}",
          f"grid/obc_x2={bcs[3]}"]
    for flux in ("roe", "hlle"):
        say(f"[gpu] tiny/thin blast {nx1}x{nx2} bcs {bcs} {flux}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, flux=flux, variant="exact")
        ora, _, _ = oracle_setup("blast", ov, flux=flux)
        compare_exact(gpu, ora, f"{nx1}x{nx2} {bcs} {flux} initial state")
        gpu.set_limits(tlim, 12)
        for s in range(12):
            if ora.clock()[1] >= tlim:
                break
            gpu.step()
            ora.step(tlim)
            compare_exact(gpu, ora, f"{nx1}x{nx2} {bcs} {flux} after {s + 1} steps")
        gpu.close()


@pytest.mark.parametrize("extra", [
    ["problem/gamma=1.4"], ["problem/gamma=2.0"], ["problem/gamma=1.1"],
    ["problem/b0=0.0"], ["problem/angle=0"], ["problem/angle=90"],
    ["problem/b0=0.0", "problem/prat=1.0"], ["problem/pamb=1e-6"]],
    ids=["gamma1.4", "gamma2", "gamma1.1", "nofield", "field_x1", "field_x2", "uniform",
         "pamb1e-6"])
def test_exact_parameter_edges_match_oracle(extra):
    """Blast 40x56 away from its default parameters: other adiabatic indices
    (gamma = 2 zeroes Gamma - 2 in the Roe eigensystem), no magnetic field
    (the eigensystems' bt = 0 and degenerate-speed branches), the field along
    x1 or x2 (bt = 0 in one sweep), a uniform state, and a near-vacuum ambient
    (the pressure floor). Both fluxes, stepped as the main loop does until
    tlim or 12 steps."""
    pkg = load_pkg()
    ov = ["grid/Nx1=40", "grid/Nx2=56"] + extra
    for flux in ("roe", "hlle"):
        say(f"[gpu] parameter edge {extra} {flux}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, flux=flux, variant="exact")
        ora, _, _ = oracle_setup("blast", ov, flux=flux)
        gpu.set_limits(tlim, 12)
        for s in range(12):
            if ora.clock()[1] >= tlim:
                break
            gpu.step()
            ora.step(tlim)
            compare_exact(gpu, ora, f"{extra} {flux} after {s + 1} steps")
        gpu.close()


@pytest.mark.parametrize("problem,extra", [
    ("field_loop", ["problem/vflow=0.0"]),
    ("field_loop", ["problem/vflow=0.0", "problem/amp=0.0"]),
    ("cpaw2d", ["grid/Nx1=32", "grid/Nx2=32", "problem/v_par=0.0"])],
    ids=["static_loop", "static_uniform", "cpaw2d_standing"])
def test_exact_zero_flow_matches_oracle(problem, extra):
    """Zero mass flux through every face at the start: a static field loop
    (the corner emf's third, averaging, upwind case everywhere, and the
    characteristic tracing's ev == 0), a static uniform state, and the
    standing circularly polarised Alfven wave."""
    pkg = load_pkg()
    for flux in ("roe", "hlle"):
        say(f"[gpu] zero flow {problem} {extra} {flux}")
        gpu, par, tlim, _ = pkg.setup_problem(problem, extra, flux=flux, variant="exact")
        ora, _, _ = oracle_setup(problem, extra, flux=flux)
        gpu.set_limits(tlim, 12)
        for s in range(12):
            gpu.step()
            ora.step(tlim)
            compare_exact(gpu, ora, f"{problem} {extra} {flux} after {s + 1} steps")
        gpu.close()


def test_batch_with_thin_grids_matches_oracle():
    """One batch of blasts of 2x2 (reflecting), 64x40 (periodic), 3x5 (mixed)
    and 40x4 (outflow, HLLE): the batch's in-place ghost fill carries every
    instance's ghosts when one of them reads stale ones, and each instance
    runs to its own tlim. Every instance equals its own oracle run."""
    pkg = load_pkg()
    cases = [(2, 2, (1, 1, 1, 1), "roe"), (64, 40, (4, 4, 4, 4), "roe"),
             (3, 5, (1, 2, 4, 1), "roe"), (40, 4, (2, 2, 2, 2), "hlle")]
    descs, params, oras = [], [], []
    for nx1, nx2, bcs, flux in cases:
        ov = [f"grid/Nx1={nx1}", f"grid/Nx2={nx2}", "problem/radius=0.2",
              f"grid/ibc_x1={bcs[0]}", f"grid/obc_x1={bcs[1]}", f"grid/ibc_x2={bcs[2]}",
              f"grid/obc_x2={bcs[3]}"]
        par = pkg.read_athinput(os.path.join(INPUTS, "athinput.blast"), ov)
        descs.append(pkg.desc_from_athinput(par, flux=flux))
        params.append(par)
        oras.append(oracle_setup("blast", ov, flux=flux))
    say(f"[gpu] batch of {len(cases)} blasts with thin grids")
    g = pkg.Grid(descs, variant="exact")
    for k, par in enumerate(params):
        pr = par["problem"]
        g.pgen_blast(k, float(pr["radius"]), float(pr["pamb"]), float(pr["prat"]),
                     float(pr["b0"]), float(pr["angle"]))
    g.set_bvals()
    g.new_dt()
    tlim = oras[0][1]
    g.set_limits(tlim, 12)
    for s in range(12):
        g.step()
        for k, (ora, _, _) in enumerate(oras):
            if ora.clock()[1] < tlim:
                ora.step(tlim)
    for k, (ora, _, _) in enumerate(oras):
        for name, x, y in zip(FIELDS, g.download(k), ora.get()):
            assert np.array_equal(x, y), f"instance {k} {cases[k]}: {name}"
        assert g.clock(k) == ora.clock(), f"instance {k} clock"
    g.close()


def _random_configs(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx1 = int(rng.choice([2, 3, 4, 5, 7, 16, 33, 64, 96]))
        nx2 = int(rng.choice([2, 3, 4, 6, 9, 17, 40, 65, 96]))
        bcs = [int(b) for b in rng.choice([1, 2, 4], size=4)]
        ov = [f"grid/Nx1={nx1}", f"grid/Nx2={nx2}",
              f"grid/ibc_x1={bcs[0]}", f"grid/obc_x1={bcs[1]}", f"grid/ibc_x2={bcs[2]}",
              f"grid/obc_x2={bcs[3]}",
              f"problem/radius={rng.uniform(0.05, 0.3):.3f}",
              f"problem/prat={rng.choice([1.0, 3.0, 10.0, 100.0])}",
              f"problem/b0={rng.choice([0.0, 0.3, 1.0, 3.0])}",
              f"problem/angle={rng.choice([0.0, 30.0, 45.0, 90.0])}",
              f"problem/gamma={rng.choice([1.4, 5.0 / 3.0, 2.0])}"]
        out.append((ov, str(rng.choice(["roe", "hlle"]))))
    return out


@pytest.mark.parametrize("k", range(4))
def test_exact_random_configurations_match_oracle(k):
    """A seeded sweep of ordinary blast configurations: sizes from 2 to 96 per
    side, every combination of boundary kinds, either flux, and random
    radius, pressure ratio, field strength and angle, adiabatic index; each
    stepped as the main loop does (until tlim or 8 steps) and compared with
    the oracle value for value after every step."""
    pkg = load_pkg()
    for ov, flux in _random_configs(1000 + k, 12):
        say(f"[gpu] random config {flux} {ov}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, flux=flux, variant="exact")
        ora, _, _ = oracle_setup("blast", ov, flux=flux)
        compare_exact(gpu, ora, f"{flux} {ov} initial state")
        gpu.set_limits(tlim, 8)
        for s in range(8):
            if ora.clock()[1] >= tlim:
                break
            gpu.step()
            ora.step(tlim)
            compare_exact(gpu, ora, f"{flux} {ov} after {s + 1} steps")
        gpu.close()


@pytest.mark.parametrize("k", range(8))
def test_exact_random_configurations_random_launch_shapes(k):
    """The random blast configurations again, each under a launch shape drawn
    at random from everything the autotuner can pick (x2 form and marching
    depth, persistent occupancy, tile order, corrector and final tile heights,
    strips, fusion depth): the kernel forms combined, not one at a time."""
    pkg = load_pkg()
    rng = np.random.default_rng(3000 + k)
    for ov, flux in _random_configs(2000 + k, 10):
        shape = dict(kind=int(rng.integers(0, 2)), seg=int(rng.choice([1, 2, 3, 5, 8, 16, 24])),
                     persist=int(rng.choice([0, 1, 4, 32])), xcd=int(rng.integers(0, 2)),
                     rows=int(rng.choice([1, 2, 4])), final=int(rng.choice([4, 8, 16])),
                     strip=int(rng.choice([0, 2, 3, 8])), fusion=int(rng.integers(0, 3)))
        say(f"[gpu] random config {flux} {ov} shape {shape}")
        gpu, par, tlim, _ = pkg.setup_problem("blast", ov, flux=flux, variant="exact")
        gpu.set_x2_kind(shape["kind"])
        gpu.set_shape(shape["seg"], shape["persist"])
        gpu.set_tile_order(shape["xcd"])
        gpu.set_corr_rows(shape["rows"])
        gpu.set_final_rows(shape["final"])
        gpu.set_strip(shape["strip"])
        gpu.set_fusion(shape["fusion"])
        ora, _, _ = oracle_setup("blast", ov, flux=flux)
        gpu.set_limits(tlim, 8)
        for s in range(8):
            if ora.clock()[1] >= tlim:
                break
            gpu.step()
            ora.step(tlim)
            compare_exact(gpu, ora, f"{flux} {ov} shape {shape} after {s + 1} steps")
        gpu.close()


def _random_smooth_configs(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        nx1 = int(rng.choice([2, 3, 5, 8, 24, 48, 80]))
        nx2 = int(rng.choice([2, 4, 7, 12, 32, 64]))
        bcs = [int(b) for b in rng.choice([1, 2, 4], size=4)]
        ov = [f"grid/Nx1={nx1}", f"grid/Nx2={nx2}",
              f"grid/ibc_x1={bcs[0]}", f"grid/obc_x1={bcs[1]}", f"grid/ibc_x2={bcs[2]}",
              f"grid/obc_x2={bcs[3]}"]
        if m % 2 == 0:
            problem = "field_loop"
            ov += [f"problem/rad={rng.uniform(0.1, 0.45):.3f}",
                   f"problem/amp={rng.choice([1e-3, 1e-1, 1.0])}",
                   f"problem/vflow={rng.choice([0.0, 0.5, 1.0, 3.0])}"]
        else:
            problem = "cpaw2d"
            ov += [f"problem/angle={rng.choice([0.0, 26.6, 45.0, 63.4, 90.0])}",
                   f"problem/b_perp={rng.choice([0.01, 0.1, 0.5])}",
                   f"problem/pres={rng.choice([0.01, 0.1, 1.0])}",
                   f"problem/v_par={rng.choice([0.0, 1.0])}"]
        out.append((problem, ov, str(rng.choice(["roe", "hlle"]))))
    return out


@pytest.mark.parametrize("k", range(3))
def test_exact_random_smooth_configurations_match_oracle(k):
    """The same for the field loop (radius, amplitude, flow speed including
    zero) and the circularly polarised Alfven wave (angle including the grid
    axes, amplitude, pressure, standing or travelling)."""
    pkg = load_pkg()
    for problem, ov, flux in _random_smooth_configs(2000 + k, 10):
        say(f"[gpu] random config {problem} {flux} {ov}")
        gpu, par, tlim, _ = pkg.setup_problem(problem, ov, flux=flux, variant="exact")
        ora, _, _ = oracle_setup(problem, ov, flux=flux)
        compare_exact(gpu, ora, f"{problem} {flux} {ov} initial state")
        gpu.set_limits(tlim, 8)
        for s in range(8):
            if ora.clock()[1] >= tlim:
                break
            gpu.step()
            ora.step(tlim)
            compare_exact(gpu, ora, f"{problem} {flux} {ov} after {s + 1} steps")
        gpu.close()


@pytest.mark.parametrize("k", range(4))
def test_batch_of_random_configurations_matches_oracle(k):
    """Twelve of the random blast configurations in one batch (one tile
    work-list, per-instance clocks, both Riemann solvers launched): every
    instance equals its own oracle run, clock included, after 8 steps. The
    first batch runs at the default launch shape, the others at one drawn at
    random from everything the autotuner can pick."""
    pkg = load_pkg()
    cfgs = _random_configs(3000 + k, 12)
    descs, params, oras = [], [], []
    for ov, flux in cfgs:
        par = pkg.read_athinput(os.path.join(INPUTS, "athinput.blast"), ov)
        descs.append(pkg.desc_from_athinput(par, flux=flux))
        params.append(par["problem"])
        oras.append(oracle_setup("blast", ov, flux=flux))
    say(f"[gpu] batch of {len(cfgs)} random blast configurations, seed {3000 + k}")
    g = pkg.Grid(descs, variant="exact")
    if k:
        rng = np.random.default_rng(4000 + k)
        g.set_x2_kind(int(rng.integers(0, 2)))
        g.set_shape(int(rng.choice([1, 3, 8, 16, 24])), int(rng.choice([0, 2, 8, 32])))
        g.set_tile_order(int(rng.integers(0, 2)))
        g.set_corr_rows(int(rng.choice([1, 2, 4])))
        g.set_final_rows(int(rng.choice([4, 8, 16])))
        g.set_strip(int(rng.choice([0, 3, 8])))
        g.set_fusion(int(rng.integers(0, 3)))
        say(f"[gpu] batch shape {g.launch_shape()} form {g.x2_kind()} fusion {g.fusion()} "
            f"strip {g.strip()} final rows {g.final_rows()}")
    for k, pr in enumerate(params):
        g.pgen_blast(k, pkg._atof(pr["radius"]), pkg._atof(pr["pamb"]), pkg._atof(pr["prat"]),
                     pkg._atof(pr["b0"]), pkg._atof(pr["angle"]))
    g.set_bvals()
    g.new_dt()
    tlim = oras[0][1]
    g.set_limits(tlim, 8)
    for s in range(8):
        g.step()
        for ora, _, _ in oras:
            if ora.clock()[1] < tlim:
                ora.step(tlim)
    for k, (ora, _, _) in enumerate(oras):
        for name, x, y in zip(FIELDS, g.download(k), ora.get()):
            assert np.array_equal(x, y), f"instance {k} {cfgs[k]}: {name}"
        assert g.clock(k) == ora.clock(), f"instance {k} clock"
    g.close()


def test_upload_download_round_trip_and_restep():
    """athmi_upload / athmi_download stage the caller's arrays in the idle
    ping-pong buffer and convert on the device: a random padded state comes
    back bit for bit, and a state re-uploaded mid-run (its staging left in the
    buffer the next step writes) keeps stepping identically to the oracle."""
    pkg = load_pkg()
    ov = ["grid/Nx1=40", "grid/Nx2=56"]
    say("[gpu] upload/download round trip, blast 40x56")
    gpu, par, tlim, _ = pkg.setup_problem("blast", ov, variant="exact")
    ora, _, _ = oracle_setup("blast", ov)
    rng = np.random.default_rng(5)
    rand = [rng.normal(size=a.shape) for a in gpu.download(0)]
    gpu.upload(0, *rand)
    for name, x, y in zip(FIELDS, gpu.download(0), rand):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), f"round trip: {name}"
    gpu.upload(0, *ora.get())
    compare_exact(gpu, ora, "re-uploaded initial state")
    gpu.set_limits(tlim, -1)
    for s in range(6):
        if s == 3:
            gpu.upload(0, *gpu.download(0))
        gpu.step()
        ora.step(tlim)
        compare_exact(gpu, ora, f"after {s + 1} steps")
    gpu.close()


@pytest.mark.parametrize("ov", [["grid/Nx1=40", "grid/Nx2=56"],
                                ["grid/Nx1=3", "grid/Nx2=2", "problem/radius=0.2"] +
                                [f"grid/{b}=1" for b in ("ibc_x1", "obc_x1", "ibc_x2", "obc_x2")]],
                         ids=["blast40x56", "blast3x2_reflect"])
def test_exact_routines_one_by_one(ov):
    """integrate_2d, new_dt and set_bvals as separate calls (the drop-in
    granularity of the reference's main loop); on a reflecting 3 x 2 grid the
    in-place set_bvals reads outer ghosts as the previous fill left them."""
    pkg = load_pkg()
    say(f"[gpu] exact routine-by-routine blast {ov}")
    gpu, par, tlim, _ = pkg.setup_problem("blast", ov, variant="exact")
    ora, _, _ = oracle_setup("blast", ov)
    for s in range(5):
        gpu.integrate()
        ora.integrate()
        compare_exact(gpu, ora, f"integrate #{s}")
        dt, t, n = gpu.clock(0)
        gpu.set_clock(0, dt, t + dt, n + 1)
        ora.set_time(dt, t + dt, n + 1)
        gpu.new_dt()
        ora.new_dt()
        compare_exact(gpu, ora, f"new_dt #{s}")
        gpu.set_bvals()
        ora.set_bvals()
        compare_exact(gpu, ora, f"set_bvals #{s}")
    gpu.close()


@pytest.mark.parametrize("problem,overrides,nsteps", [
    ("field_loop", [], 100),
    ("cpaw2d", ["grid/Nx1=64", "grid/Nx2=64"], 100),
])
@pytest.mark.parametrize("flux", ["roe", "hlle"])
def test_fast_variant_l1_within_1e12(problem, overrides, nsteps, flux):
    """The FMA variant on smooth flows. (On the blast wave FMA rounding flips
    discontinuous branches -- the Roe->HLLE fallback test and the ev = 0
    switch of the characteristic tracing -- and the difference grows far past
    1e-12 within 100 steps, so the exact variant is the parity path.)"""
    pkg = load_pkg()
    say(f"[gpu] fast {problem} {flux} {overrides} steps={nsteps}")
    gpu, par, tlim, _ = pkg.setup_problem(problem, overrides, flux=flux, variant="fast")
    ora, _, _ = oracle_setup(problem, overrides, flux=flux)
    gpu.set_limits(tlim, nsteps)
    gpu.run()
    ora.run(tlim, nsteps)
    a = gpu.download(0)
    b = ora.get()
    nx1, nx2 = gpu.descs[0]["nx1"], gpu.descs[0]["nx2"]
    worst = 0.0
    for q in range(8):
        worst = max(worst, l1_rel(interior(a[0][..., q], nx1, nx2), interior(b[0][..., q], nx1, nx2)))
    for k in (1, 2, 3):
        worst = max(worst, l1_rel(interior(a[k], nx1, nx2), interior(b[k], nx1, nx2)))
    say(f"    worst L1-relative difference over fields: {worst:.3e}")
    assert worst <= 1e-12
    gpu.close()


def divb_scale(gpu):
    gas, b1, b2, b3 = gpu.download(0)
    d = gpu.descs[0]
    dx = min((d["x1max"] - d["x1min"]) / d["nx1"], (d["x2max"] - d["x2min"]) / d["nx2"])
    bmax = max(np.abs(b1).max(), np.abs(b2).max())
    return bmax / dx


@pytest.mark.parametrize("variant", ["exact", "fast"])
def test_divb_stays_zero_every_step(variant):
    """CT keeps div B at roundoff every step. The field loop (|B| ~ 1e-3)
    meets max|div B| <= 1e-13 outright; for the blast (|B| ~ 1, |B|/dx ~ 1e2)
    a few ulp of B over dx already exceed 1e-13 in the reference's own
    arithmetic (the exact variant reproduces it bit for bit), so the bound is
    stated relative to max|B|/dx there."""
    pkg = load_pkg()
    say(f"[gpu] div B {variant} field_loop + blast")
    for problem, ov in (("field_loop", []), ("blast", ["grid/Nx1=128", "grid/Nx2=192"])):
        gpu, par, tlim, _ = pkg.setup_problem(problem, ov, variant=variant)
        gpu.set_limits(tlim, 60)
        scale = divb_scale(gpu)
        for s in range(61):
            dv = gpu.divb_max(0)
            if problem == "field_loop":
                assert dv <= 1e-13, f"{problem} step {s}: max|div B| = {dv}"
            assert dv / scale <= 1e-14, f"{problem} step {s}: max|div B| dx/|B| = {dv / scale}"
            gpu.step()
        gpu.close()


@pytest.mark.parametrize("flux", ["roe", "hlle"])
def test_gpu_cpaw2d_converges_at_second_order(flux):
    """cpaw2d (the inclined circularly polarised Alfven wave of cpaw2d.c:56-185,
    an exact travelling solution) run for one wave period on the GPU at
    N = 64 .. 512 (Lx = 2 Ly): the L1 error against the initial state
    (cpaw2d.c:222-302) falls at about second order."""
    pkg = load_pkg()
    errs = []
    for n in (64, 128, 256, 512):
        print(f"[gpu] cpaw2d {flux} {n}x{n} one period", flush=True)
        g, par, tlim, _ = pkg.setup_problem("cpaw2d", [f"grid/Nx1={n}", f"grid/Nx2={n}"],
                                            flux=flux)
        g.set_limits(tlim, -1)
        g.run()
        rms = g.cpaw2d_l1(0)[0]
        errs.append(rms)
        g.close()
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    overall = np.log2(errs[0] / errs[-1]) / (len(errs) - 1)
    print(f"[gpu] cpaw2d {flux} RMS L1 errors {errs} orders {orders} overall {overall}",
          flush=True)
    # the limiter clips the wave's extrema, so the local order dips below 2 at
    # fine resolution (HLLE more than Roe); the fitted order stays near 2
    assert np.all(orders > 1.7) and overall > 1.85, (orders, overall)
