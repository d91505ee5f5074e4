"""Parity of the HIP path (through the C ABI) with the oracle and the golden fixtures, on an MI355X.

Bar: CSR pattern and slot map bit-exact; energy, gradient and Hessian values within 1e-12 relative max-norm
(|a - b|_inf / |b|_inf per output array) of the oracle, for Exact and Projected Hessians.
"""
import hashlib
import math
from pathlib import Path

import numpy as np
import pytest

from litepgo_amd import VolumetricFemEnergy, kuhn_grid, lame_parameters, workloads
from litepgo_amd import _native

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert _native.LIB_PATH.exists(), "liblitepgo_fem.so is not built"
    return torch


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.abs(b).max() if b.size else 0.0
    if scale == 0.0:
        return float(np.abs(a).max()) if a.size else 0.0
    return float(np.abs(a - b).max() / scale)


def check_against_oracle(dev, ref, x, kinds=("exact", "projected"), theta=None, tol=TOL):
    for kind in kinds:
        print(f"  kind={kind}", flush=True)
        v, g, H = dev.eval_values(x, True, kind)
        vr, gr, Hr = ref.eval(x, True, kind, theta=theta)
        assert abs(v - vr) <= tol * max(abs(vr), 1e-300), (kind, v, vr)
        assert rel(g, gr) <= tol, (kind, rel(g, gr))
        assert rel(H, Hr) <= tol, (kind, rel(H, Hr))


def make(oracle_mod, X, tets, theta):
    return VolumetricFemEnergy(len(X), tets, X, theta), oracle_mod.OracleEnergy(len(X), tets, X, theta)


def test_two_tet_fixture(oracle_mod):
    d = np.load(GOLD / "two_tet.npz")
    dev = VolumetricFemEnergy(5, d["tets"], d["X"], d["theta"])
    assert np.array_equal(dev.pattern()[0], d["outer"]) and np.array_equal(dev.pattern()[1], d["inner"])
    assert np.array_equal(dev.slot_map(), d["offsets"])
    for s in range(3):
        v, g, H = dev.eval_values(d["x"][s], True, "exact")
        _, _, Hp = dev.eval_values(d["x"][s], False, "projected")
        assert abs(v - d["value_exact"][s]) <= TOL * max(abs(d["value_exact"][s]), 1e-300)
        assert rel(g, d["grad"][s]) <= TOL
        assert rel(H, d["H_exact"][s]) <= TOL
        assert rel(Hp, d["H_projected"][s]) <= TOL


def test_grid2_fixture_and_oracle(oracle_mod):
    d = np.load(GOLD / "grid2.npz")
    dev, ref = make(oracle_mod, d["X"], d["tets"], d["theta"])
    for s in range(2):
        print(f"grid2 state {s}", flush=True)
        v, g, H = dev.eval_values(d["x"][s], True, "exact")
        _, _, Hp = dev.eval_values(d["x"][s], False, "projected")
        assert abs(v - d["value_exact"][s]) <= TOL * abs(d["value_exact"][s])
        assert rel(g, d["grad"][s]) <= TOL
        assert rel(H, d["H_exact"][s]) <= TOL
        assert rel(Hp, d["H_projected"][s]) <= TOL
        check_against_oracle(dev, ref, d["x"][s])
    assert dev.last_indefinite_count > 0  # the compressed state has indefinite elements


def test_plain_small_config(oracle_mod):
    """10^3 Kuhn grid, E = 1e5, nu = 0.4, u ~ U(+-0.05 h): the reference's CPU-sized configuration."""
    w = workloads.plain_small()
    dev, ref = make(oracle_mod, w.rest, w.tets, w.theta)
    check_against_oracle(dev, ref, w.u)
    d = np.load(GOLD / "grid10.npz")
    v, g, H = dev.eval_values(d["x"], True, "exact")
    _, _, Hp = dev.eval_values(d["x"], False, "projected")
    s = d["sample"]
    assert abs(v - d["value_exact"]) <= TOL * abs(d["value_exact"])
    assert rel(g, d["grad"]) <= TOL
    assert np.abs(H[s] - d["H_exact_sample"]).max() <= TOL * d["H_exact_absmax"]
    assert np.abs(Hp[s] - d["H_projected_sample"]).max() <= TOL * d["H_projected_absmax"]
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    assert sha(dev.pattern()[0]) == str(d["outer_sha"]) and sha(dev.pattern()[1]) == str(d["inner_sha"])
    assert sha(dev.slot_map()) == str(d["offsets_sha"])


def test_device_tensor_interface_matches_host_interface(gpu):
    torch = gpu
    w = workloads.plain_small(seed=3)
    dev = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    x = torch.from_numpy(w.u).cuda()
    for kind in ("exact", "projected"):
        v, g, H = dev.eval_device(x, True, kind)
        vh, gh, Hh = dev.eval_values(w.u, True, kind)
        assert v == vh
        assert np.array_equal(g.cpu().numpy(), gh) and np.array_equal(H.cpu().numpy(), Hh)
    v, g, H = dev.eval_device(x)
    assert g is None and H is None and v == dev.eval_value(w.u)


LAUNCH_CONFIGS = [(4, 1, 16, 4096), (4, 1, 8, 4096), (4, 1, 12, 4096), (4, 1, 3, 4096), (4, 2, 32, 6144),
                  (4, 2, 7, 6144), (4, 4, 64, 12288), (4, 4, 40, 8000), (4, 1, 16, 200), (4, 2, 32, 300)]


@pytest.mark.parametrize("cfg", LAUNCH_CONFIGS)
def test_launch_configurations(oracle_mod, cfg):
    """Every launch shape meets the oracle bar for Exact and Projected Hessians and gives the default shape's results
    bit for bit: the summation order of every value is fixed by the mesh (the reference's assembly is bitwise
    thread-count invariant, tests/cpp/energy/scalar/mesh_energy_gtest.cpp:316-359).  The last two configurations'
    tiny LDS budgets send every tile through the HBM-accumulation path, which must keep the same order."""
    w = workloads.plain(6, 1e5, 0.4, 0.15, seed=5)
    dev, ref = make(oracle_mod, w.rest, w.tets, w.theta)
    base = [dev.eval_values(w.u, True, k) for k in ("exact", "projected")]
    dev.set_launch(*cfg)
    assert dev.get_launch() == cfg
    check_against_oracle(dev, ref, w.u)
    for k, b in zip(("exact", "projected"), base):
        v1, g1, H1 = dev.eval_values(w.u, True, k)
        assert v1 == b[0] and np.array_equal(g1, b[1]) and np.array_equal(H1, b[2]), (cfg, k)


def test_other_lane_counts_are_rejected():
    w = workloads.plain(3, 1e5, 0.4, 0.1, seed=1)
    dev = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    dev.eval_values(w.u, True, "exact")
    assert dev.get_launch() == (4, 1, 16, 4096)
    for cfg in [(8, 1, 8, 4096), (5, 1, 12, 4096), (2, 1, 32, 6144), (4, 3, 48, 4096), (4, 1, 17, 4096)]:
        with pytest.raises(ValueError):
            dev.set_launch(*cfg)
    assert dev.get_launch() == (4, 1, 16, 4096)


def test_run_to_run_determinism():
    w = workloads.plain(12, 1e5, 0.45, 0.1, stretch=0.1, seed=9)
    dev = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    a = dev.eval_values(w.u, True, "projected")
    b = dev.eval_values(w.u, True, "projected")
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_projection_far_from_rest(oracle_mod):
    """States whose tangents are indefinite everywhere: a 0.6x compression (3 negative eigenvalues per tet: the
    projection adds a rank-3 term to the exact blocks) and 1.6x / 2x dilations (5 negative eigenvalues: the rank-4
    positive part replaces the exact blocks).  Every tet is listed, so every lane takes the rescan path too."""
    rng = np.random.default_rng(17)
    X, tets = kuhn_grid(4)
    lam, mu = lame_parameters(1e5, 0.45)
    theta = np.tile([lam, mu], (len(tets), 1))
    dev, ref = make(oracle_mod, X, tets, theta)
    for scale in (-0.4, 0.6, 1.0):
        print(f"uniform scale {1 + scale}", flush=True)
        u = (scale * X + 0.0025 * rng.uniform(-1, 1, X.shape)).reshape(-1)
        check_against_oracle(dev, ref, u)
        assert dev.last_indefinite_count == len(tets)


def regular_tets_scaled(count, scales, seed):
    """Disjoint regular tetrahedra (random size, orientation and place), each deformed by F = s R (s from `scales` in
    turn, R a random rotation).  For a regular tet the reduced element Hessian is a multiple of the tangent in a
    rotated basis, and at F = s R the tangent's spectrum is {mu + p/s^2 (x3), mu - p/s^2 (x5), mu - p/s^2 + 3 lam/s^2}:
    exactly repeated eigenvalues."""
    rng = np.random.default_rng(seed)
    base = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)

    def rotation():
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        return Q if np.linalg.det(Q) > 0 else -Q

    X, U, tets = [], [], []
    for t in range(count):
        P = base @ rotation().T * (0.02 * (1.0 + rng.uniform())) + rng.uniform(-1, 1, 3)
        c = P.mean(0)
        x = (P - c) @ (scales[t % len(scales)] * rotation()).T + c
        X.append(P)
        U.append(x - P)
        tets.append([4 * t, 4 * t + 2, 4 * t + 1, 4 * t + 3])  # (0, 2, 1, 3): positive orientation of `base`
    return np.concatenate(X), np.array(tets, dtype=np.int32), np.concatenate(U).reshape(-1)


def test_projection_with_exactly_repeated_eigenvalues(oracle_mod):
    """Element Hessians with exactly repeated eigenvalues (regular tets under F = s R, nu = 0.45): s = 0.7 gives a
    3-fold negative eigenvalue (its eigenspace is added), s = 1.5 and 2 a 5-fold negative one (the positive part,
    a 3-fold eigenvalue and a single one, replaces the exact blocks), s = 1 a stress-free rotation (semi-definite
    with a 3-fold zero eigenvalue: listed, and the clamp returns M itself).  The eigen-clamp's last inverse-iteration
    vectors inside a repeated eigenvalue are the hard case."""
    X, tets, u = regular_tets_scaled(96, (0.7, 1.5, 2.0, 1.0), seed=23)
    lam, mu = lame_parameters(1e5, 0.45)
    theta = np.tile([lam, mu], (len(tets), 1))
    dev, ref = make(oracle_mod, X, tets, theta)
    check_against_oracle(dev, ref, u)
    assert dev.last_indefinite_count == len(tets)


def test_real_mesh_from_veg(oracle_mod):
    """The reference's box-tet.veg asset (243 vertices, 827 irregular tets), read with litepgo_amd.read_veg: the
    state of the reference's own test (u = 1e-3 sin(i)) and a 30-degree twist about the long axis."""
    from litepgo_amd import read_veg

    mesh = read_veg(GOLD / "box-tet.veg")
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([lam, mu], (len(mesh.cells), 1))
    dev, ref = make(oracle_mod, mesh.positions, mesh.cells, theta)
    n = 3 * mesh.num_vertices
    check_against_oracle(dev, ref, 1e-3 * np.sin(np.arange(n)))
    X = mesh.positions
    z = (X[:, 2] - X[:, 2].min()) / np.ptp(X[:, 2])
    ang = np.deg2rad(30.0) * z
    cx, cy = X[:, 0].mean(), X[:, 1].mean()
    rx, ry = X[:, 0] - cx, X[:, 1] - cy
    u = np.zeros_like(X)
    u[:, 0] = np.cos(ang) * rx - np.sin(ang) * ry - rx
    u[:, 1] = np.sin(ang) * rx + np.cos(ang) * ry - ry
    check_against_oracle(dev, ref, u.reshape(-1))


def test_irregular_numbering_and_unused_vertex(oracle_mod):
    rng = np.random.default_rng(21)
    X, tets = kuhn_grid((3, 4, 5))
    perm_v = rng.permutation(len(X))
    inv = np.empty_like(perm_v)
    inv[perm_v] = np.arange(len(X))
    X = np.vstack([X[perm_v], [[7.0, 7.0, 7.0]]])
    tets = inv[tets][rng.permutation(len(tets))].astype(np.int32)
    theta = np.stack([rng.uniform(1e3, 1e5, len(tets)), rng.uniform(1e3, 1e5, len(tets))], 1)
    dev, ref = make(oracle_mod, X, tets, theta)
    x = rng.uniform(-0.04, 0.04, X.size)
    x[-3:] = 0.5  # the unused vertex: zero gradient, empty rows
    check_against_oracle(dev, ref, x)
    v, g, H = dev.eval_values(x, True, "exact")
    assert np.all(g[-3:] == 0.0)


def fan_mesh(n, seed):
    """n well-shaped tets sharing vertex 0 (distinct other vertices): vertex 0 has degree 3 n + 1."""
    rng = np.random.default_rng(seed)
    X = [np.zeros(3)]
    tets = []
    for _ in range(n):
        base = rng.normal(size=3)
        base /= np.linalg.norm(base)
        t1 = np.cross(base, rng.normal(size=3))
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(base, t1)
        pts = [2 * base + 0.6 * t1, 2 * base + 0.6 * t2, 2.6 * base]
        t = [0, len(X), len(X) + 1, len(X) + 2]
        X.extend(pts)
        D = np.stack([X[t[k]] - X[t[0]] for k in (1, 2, 3)], 1)
        if np.linalg.det(D) < 0:
            t[1], t[2] = t[2], t[1]
        tets.append(t)
    return np.array(X), np.array(tets, dtype=np.int32)


def test_high_degree_vertex_uses_wide_incidences(oracle_mod):
    """A fan of 120 tets around one vertex: its degree (361) exceeds the 8-bit column-position form."""
    n = 120
    X, tets = fan_mesh(n, 4)
    theta = np.tile([[1e4, 2e4]], (n, 1))
    dev, ref = make(oracle_mod, X, tets, theta)
    assert dev.max_degree == 3 * n + 1
    x = np.random.default_rng(5).uniform(-0.01, 0.01, X.size)
    check_against_oracle(dev, ref, x)
    # flattening vertex 0's tets inverts them: both sides raise the domain error
    x = np.zeros(X.size)
    x[3:] = -1.2 * X[1:].reshape(-1)
    with pytest.raises(oracle_mod.OracleError) as e:
        ref.eval(x, True, "exact")
    assert e.value.code == 2
    with pytest.raises(_native.DomainError):
        dev.eval_values(x, True, "exact")


def test_high_degree_vertex_force_fp_and_products(oracle_mod, gpu):
    """The wide-incidence fan through the vertex-gather products: the plastic force response (its gather reads the
    wide records) and the Hessian-vector product, against the oracle."""
    torch = gpu
    n = 120
    X, tets = fan_mesh(n, 4)
    theta = np.tile([[1e4, 2e4]], (n, 1))
    dev = VolumetricFemEnergy(len(X), tets, X, theta)
    assert dev.max_degree == 3 * n + 1
    rng = np.random.default_rng(6)
    u = rng.uniform(-0.01, 0.01, X.size)
    Fp = np.eye(3)[None] + 0.1 * rng.normal(size=(n, 3, 3))
    fp = np.ascontiguousarray(Fp.transpose(0, 2, 1).reshape(n, 9))
    d = rng.normal(size=(n, 9))
    fr, _, jvr = oracle_mod.force_fp_products(tets, X, theta, u, fp, dFp=d)
    x = torch.from_numpy(u).cuda()
    fpd = torch.from_numpy(fp).cuda()
    assert rel(dev.force_fp_device(x, fpd).cpu().numpy(), fr) <= TOL
    assert rel(dev.force_jvp_fp_device(x, torch.from_numpy(d).cuda(), fpd).cpu().numpy(), jvr) <= TOL
    ref = oracle_mod.OracleEnergy(len(X), tets, X, theta)
    _, _, Hr = ref.eval(u, False, "exact")
    outer, inner = ref.pattern()
    import scipy.sparse as sp

    v = rng.normal(size=u.size)
    Hv = sp.csc_matrix((Hr, inner, outer), shape=(ref.num_dofs, ref.num_dofs)) @ v
    assert rel(dev.hvp_device(x, torch.from_numpy(v).cuda()).cpu().numpy(), Hv) <= TOL


@pytest.mark.parametrize("num_vertices", [0, 4])
def test_empty_mesh(gpu, num_vertices):
    """No tets (the reference's zero-element meshes): E = 0, a zero gradient, an empty Hessian, zero products."""
    torch = gpu
    X = np.eye(4, 3)[:num_vertices]
    dev = VolumetricFemEnergy(num_vertices, np.zeros((0, 4), dtype=np.int32), X, np.zeros((0, 2)))
    assert dev.nnz == 0 and dev.num_dofs == 3 * num_vertices
    x = np.full(dev.num_dofs, 0.1)
    for kind in ("exact", "projected"):
        print(f"  V={num_vertices} kind={kind}", flush=True)
        v, g, H = dev.eval_values(x, True, kind)
        assert v == 0.0 and g.shape == (dev.num_dofs,) and np.all(g == 0.0) and H.size == 0
    xd = torch.from_numpy(x).cuda()
    fp = torch.zeros(0, dtype=torch.float64, device="cuda")
    assert torch.all(dev.force_fp_device(xd, fp) == 0.0) and dev.force_fp_device(xd, fp).numel() == dev.num_dofs
    assert dev.force_vjp_fp_device(xd, xd, fp).shape == (0, 9)
    assert dev.max_feasible_step(x, x) == 1.0


def test_error_semantics_and_recovery():
    w = workloads.plain(4, 1e5, 0.4, 0.05, seed=1)
    dev = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    good = dev.eval_values(w.u, True, "exact")
    x = w.u.copy()
    x[5] = math.nan
    with pytest.raises(ValueError, match="finite"):
        dev.eval_values(x, True, "exact")
    inverted = -2.0 * w.rest.reshape(-1)  # F = -I everywhere
    with pytest.raises(_native.DomainError):
        dev.eval_values(inverted, True, "exact")
    huge = w.u.copy()
    huge[0] = 1e155  # finite input whose density overflows
    with pytest.raises(_native.DomainError):
        dev.eval_values(huge)
    bad = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, np.vstack([w.theta[:-1], [[-1.0, 1.0]]]))
    with pytest.raises(ValueError, match="lambda"):
        bad.eval_values(w.u)
    with pytest.raises(ValueError, match="HessianKind|Hessian kind"):
        dev.eval_values(w.u, False, 2)
    again = dev.eval_values(w.u, True, "exact")  # the object stays usable
    assert again[0] == good[0] and np.array_equal(again[1], good[1]) and np.array_equal(again[2], good[2])


def test_theta_override_and_derivatives(oracle_mod, gpu):
    torch = gpu
    w, E, nu = workloads.differentiate(n=6, seed=2)
    dev, ref = make(oracle_mod, w.rest, w.tets, w.theta)
    x = torch.from_numpy(w.u).cuda()
    theta2 = w.theta * np.random.default_rng(0).uniform(0.5, 1.5, w.theta.shape)
    t2 = torch.from_numpy(theta2).cuda()
    v, g, H = dev.eval_device(x, True, "exact", theta=t2)
    vr, gr, Hr = ref.eval(w.u, True, "exact", theta=theta2)
    assert abs(v - vr) <= TOL * abs(vr) and rel(g.cpu().numpy(), gr) <= TOL and rel(H.cpu().numpy(), Hr) <= TOL
    # dE/dtheta
    assert rel(dev.dE_dtheta_device(x).cpu().numpy(), ref.dE_dtheta(w.u)) <= TOL
    # force VJP / JVP in theta
    rng = np.random.default_rng(1)
    wv = rng.normal(size=w.u.size)
    assert rel(dev.force_vjp_theta_device(x, torch.from_numpy(wv).cuda()).cpu().numpy(),
               ref.force_theta_vjp(w.u, wv)) <= TOL
    dth = rng.normal(size=w.theta.shape)
    assert rel(dev.force_jvp_theta_device(x, torch.from_numpy(dth).cuda()).cpu().numpy(),
               ref.force_theta_jvp(w.u, dth)) <= TOL
    # Hessian-vector product against the oracle's assembled Exact Hessian
    import scipy.sparse as sp

    _, _, Hr = ref.eval(w.u, False, "exact")
    outer, inner = ref.pattern()
    Hm = sp.csc_matrix((Hr, inner, outer), shape=(ref.num_dofs, ref.num_dofs))
    vv = rng.normal(size=w.u.size)
    assert rel(dev.hvp_device(x, torch.from_numpy(vv).cuda()).cpu().numpy(), Hm @ vv) <= TOL


def test_plain_large_config_full_size(oracle_mod):
    """64^3 Kuhn grid (1,572,864 tets), E = 1e5, nu = 0.45, 20% stretch + U(+-0.02 h): full-size parity with the
    oracle plus size-independent invariants (translation invariance, zero net force)."""
    w = workloads.plain_large()
    print("building device energy", flush=True)
    dev = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    print("building oracle", flush=True)
    ref = oracle_mod.OracleEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    assert dev.nnz == ref.nnz == 36168777
    r, c = dev.pattern()
    ro, co = ref.pattern()
    assert np.array_equal(r, ro) and np.array_equal(c, co)
    print("comparing values", flush=True)
    check_against_oracle(dev, ref, w.u)
    v, g, H = dev.eval_values(w.u, True, "exact")
    # a rigid translation is in the null space of every element Hessian; the net force vanishes
    import scipy.sparse as sp

    Hm = sp.csr_matrix((H, c, r), shape=(dev.num_dofs, dev.num_dofs))
    for axis in range(3):
        t = np.zeros(dev.num_dofs)
        t[axis::3] = 1.0
        assert np.abs(Hm @ t).max() <= 1e-12 * np.abs(H).max() * 30
    assert abs(g[0::3].sum()) <= 1e-10 * np.abs(g).sum()


def test_projection_store_grows_when_more_tets_are_listed(oracle_mod):
    """The low-rank projection store starts at max(4096, T / 16) entries.  A compressed 10^3 grid lists all 6,000
    tets: the first evaluation overflows the store, the library grows it and evaluates again, and the result meets
    the oracle bar; graph mode then captures the grown store, and a milder state reuses it."""
    rng = np.random.default_rng(23)
    X, tets = kuhn_grid(10)
    lam, mu = lame_parameters(1e5, 0.45)
    theta = np.tile([lam, mu], (len(tets), 1))
    dev, ref = make(oracle_mod, X, tets, theta)
    u = (-0.4 * X + 0.002 * rng.uniform(-1, 1, X.shape)).reshape(-1)
    check_against_oracle(dev, ref, u)
    assert dev.last_indefinite_count == len(tets)
    dev.set_graph(True)
    first = dev.eval_values(u, True, "projected")
    for _ in range(3):
        again = dev.eval_values(u, True, "projected")
        assert again[0] == first[0] and np.array_equal(again[1], first[1]) and np.array_equal(again[2], first[2])
    dev.set_graph(False)
    u2 = (0.002 * rng.uniform(-1, 1, X.shape)).reshape(-1)
    check_against_oracle(dev, ref, u2)


def test_graph_replays_report_energy_status_and_count():
    """In graph mode a single-mesh evaluation hands energy, status word and list count to the host from its last
    kernel.  Replays of one captured graph with new contents in the same buffers: the energy and the indefinite count
    match an eager evaluation of the same state bit for bit, a non-finite x or an inverted element raises the
    reference's error class, and the handle recovers."""
    import torch

    w = workloads.plain(6, 1e5, 0.45, 0.05, stretch=0.2, seed=4)
    dev = VolumetricFemEnergy(w.num_vertices, w.tets, w.rest, w.theta)
    states = [w.u, w.u * 1.5, 0.5 * w.u]
    eager = []
    for u in states:
        v, g, H = dev.eval_device(torch.from_numpy(u).cuda(), True, "projected")
        eager.append((v, dev.last_indefinite_count, g.cpu().numpy(), H.cpu().numpy()))
    assert any(e[1] > 0 for e in eager)
    x = torch.from_numpy(w.u.copy()).cuda()
    g = torch.empty(dev.num_dofs, dtype=torch.float64, device="cuda")
    H = torch.empty(dev.nnz, dtype=torch.float64, device="cuda")
    dev.set_graph(True)
    try:
        for _ in range(2):  # eager, then captured
            dev.eval_device(x, True, "projected", out_g=g, out_H=H)
        for rep in range(3):
            for (v, count, ge, He), u in zip(eager, states):
                x.copy_(torch.from_numpy(u))
                got = dev.eval_device(x, True, "projected", out_g=g, out_H=H)[0]
                assert got == v and dev.last_indefinite_count == count, (rep, got, v)
                assert np.array_equal(g.cpu().numpy(), ge) and np.array_equal(H.cpu().numpy(), He)
                torch.cuda.synchronize()
        bad = w.u.copy()
        bad[7] = math.nan
        x.copy_(torch.from_numpy(bad))
        with pytest.raises(ValueError, match="finite"):
            dev.eval_device(x, True, "projected", out_g=g, out_H=H)
        x.copy_(torch.from_numpy(-2.0 * w.rest.reshape(-1)))
        with pytest.raises(_native.DomainError):
            dev.eval_device(x, True, "projected", out_g=g, out_H=H)
        x.copy_(torch.from_numpy(w.u))
        assert dev.eval_device(x, True, "projected", out_g=g, out_H=H)[0] == eager[0][0]
    finally:
        dev.set_graph(False)
