"""Pins of the FP64 oracle against what the mathematics fixes: textbook Fresnel, the HG and
biased pdfs and their inverse CDFs, the importance-sampling identities of equations (8)
and (9), brute-force ray-tetrahedron exits and the glass-slab known answer."""
import math

import numpy as np
import pytest
from scipy import integrate

import octmps as O
import oracle_py as OR


# ------------------------------------------------------------------ Fresnel
def fresnel_textbook(ni, nt, ci):
    """Unpolarised Fresnel reflectance, 1/2 (Rs + Rp), from the angles."""
    ti = math.acos(ci)
    st = ni / nt * math.sin(ti)
    if st >= 1.0:
        return 1.0
    tt = math.asin(st)
    if ti == 0.0:
        return ((ni - nt) / (ni + nt)) ** 2
    rs = math.sin(ti - tt) ** 2 / math.sin(ti + tt) ** 2
    rp = math.tan(ti - tt) ** 2 / math.tan(ti + tt) ** 2
    return 0.5 * (rs + rp)


def test_fresnel_matches_textbook():
    rng = np.random.default_rng(0)
    for _ in range(3000):
        ni, nt = rng.uniform(1.0, 1.7, 2)
        ci = rng.uniform(1e-5, 1.0)
        ref = fresnel_textbook(ni, nt, ci)
        got = OR.fresnel(ni, nt, ci)
        if ci < 1e-6:
            assert got == 1.0
        else:
            assert abs(got - ref) < 1e-12 + 1e-9 * ref, (ni, nt, ci, got, ref)


def test_fresnel_special_cases():
    assert abs(OR.fresnel(1.4, 1.0, 1.0) - (0.4 / 2.4) ** 2) < 1e-15          # normal incidence
    assert OR.fresnel(1.5, 1.0, 0.5) == 1.0                                   # beyond the critical angle
    assert OR.fresnel(1.33, 1.33, 0.3) < 1e-25                                # matched indices
    assert OR.fresnel(1.0, 1.5, 5e-7) == 1.0                                  # grazing (COSNINETYDEG)


# ------------------------------------------------------------------ scattering pdfs
def hg_pdf(mu, g):
    return (1 - g * g) / (2 * (1 + g * g - 2 * g * mu) ** 1.5)


def bias_pdf(mu, a):
    s = math.sqrt(1 + a * a)
    return a * (1 - a) * s / ((s - 1 + a) * (1 + a * a - 2 * a * mu) ** 1.5)


@pytest.mark.parametrize("g", [0.9, 0.5, 0.95, -0.3])
def test_hg_sampler_is_inverse_cdf(g):
    assert abs(integrate.quad(hg_pdf, -1, 1, args=(g,))[0] - 1) < 1e-10
    for u in np.linspace(0.01, 1.0, 23):
        mu = OR.spin_theta(g, u)
        cdf = integrate.quad(hg_pdf, -1, mu, args=(g,), epsabs=1e-13)[0]
        assert abs(cdf - u) < 1e-8, (u, mu, cdf)


@pytest.mark.parametrize("a", [0.925, 0.5, 0.1])
def test_biased_sampler_is_inverse_cdf_on_0_1(a):
    assert abs(integrate.quad(bias_pdf, 0, 1, args=(a,))[0] - 1) < 1e-10
    for u in np.linspace(0.01, 1.0, 23):
        mu = OR.spin_theta_bias(a, u)
        assert 0.0 <= mu <= 1.0
        cdf = integrate.quad(bias_pdf, 0, mu, args=(a,), epsabs=1e-13)[0]
        assert abs(cdf - u) < 1e-8, (u, mu, cdf)


# ------------------------------------------------------------------ SpinBias identities
def hemisphere_hg_mass(u0, v, g):
    """P_HG(omega . v >= 0) for HG scattering around u0, by quadrature over the sphere."""
    u0 = np.asarray(u0, float) / np.linalg.norm(u0)
    v = np.asarray(v, float) / np.linalg.norm(v)
    # frame around u0
    e1 = np.cross(u0, [1.0, 0, 0] if abs(u0[0]) < 0.9 else [0, 1.0, 0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u0, e1)

    def f(phi, mu):
        s = math.sqrt(max(0.0, 1 - mu * mu))
        w = mu * u0 + s * (math.cos(phi) * e1 + math.sin(phi) * e2)
        return hg_pdf(mu, g) / (2 * math.pi) if w @ v >= 0 else 0.0

    val, _ = integrate.dblquad(f, -1, 1, 0, 2 * math.pi, epsabs=1e-7)
    return val


def spinbias_params(type_bias=37):
    return O.make_params([(1.4, 1.0, 100.0, 0.9)], num_bins=64, depth=0.1, type_bias=type_bias,
                         backward_bias_coefficient=0.925, probability_additional_bias=0.5,
                         max_collecting_angle_deg=5.0, rnd_step_size_in_tissue=0.0)


def test_first_bias_lr_is_hg_over_biased_pdf(table):
    # A forward photon inside the target depth gets its first biased scatter: F = 1, the new
    # direction is drawn around v (towards the probe) and LR *= pdf_HG(u0.u')/pdf_B(v.u')
    # (equation 8). Its mean over q is the HG mass of the hemisphere v.u' >= 0.
    p = spinbias_params()
    g, a = 0.9, 0.925
    u0 = np.array([0.3, -0.2, 0.9])
    u0 /= np.linalg.norm(u0)
    pos = np.array([0.004, 0.003, 0.05])
    state = [*pos, *u0, 0.5, 1.0, 0, 0, g, 101.0]
    n = 400000
    out = OR.spin_bias_batch(p, state, (0.0, 0.0, 0.0), 7, table, n)
    assert np.all(out[:, 4] == 1) and np.all(out[:, 7] == 1)
    v = -pos / np.linalg.norm(pos)
    w = out[:, :3]
    assert np.allclose(np.linalg.norm(w, axis=1), 1, atol=1e-12)
    # factor == pdf_HG(u0.w) / pdf_B(v.w) exactly as the two densities
    mu_u, mu_v = w @ u0, w @ v
    assert np.all(mu_v >= -1e-12)
    expect = hg_pdf(mu_u, g) / bias_pdf(np.clip(mu_v, 0, 1), a)
    assert np.allclose(out[:, 3], expect, rtol=1e-9)
    mass = hemisphere_hg_mass(u0, v, g)
    m, se = out[:, 3].mean(), out[:, 3].std() / math.sqrt(n)
    assert abs(m - mass) < 5 * se, (m, mass, se)


def sphere_integral(u0, fun, g):
    """Integral over the unit sphere of fun(w) with a frame around u0; the HG peak is
    resolved by integrating in t = sqrt(1+g^2-2g mu)."""
    u0 = np.asarray(u0, float) / np.linalg.norm(u0)
    e1 = np.cross(u0, [1.0, 0, 0] if abs(u0[0]) < 0.9 else [0, 1.0, 0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u0, e1)
    tq, tw = np.polynomial.legendre.leggauss(400)
    pq, pw = np.polynomial.legendre.leggauss(256)
    t = (1 - g) + (tq + 1) / 2 * ((1 + g) - (1 - g))
    wt = tw / 2 * 2 * g
    mu = (1 + g * g - t * t) / (2 * g)        # dmu = t dt / g
    jac = t / g
    phi = (pq + 1) * math.pi
    wphi = pw * math.pi
    s = np.sqrt(np.maximum(0, 1 - mu * mu))
    W = (mu[:, None, None] * u0 + s[:, None, None] *
         (np.cos(phi)[None, :, None] * e1 + np.sin(phi)[None, :, None] * e2))
    vals = fun(W.reshape(-1, 3)).reshape(len(mu), len(phi), -1)
    return np.einsum("ijk,i,j->k", vals, wt * jac, wphi)


def test_additional_bias_mixture_identity(table):
    # A photon already biased once, outside the collection cone, scatters from the mixture
    # p q_v + (1-p) HG_u0 with LR *= HG/(p q~ + (1-p) HG) (equation 9,
    # src/octmps_kernel.cuh:1055-1078). The reference evaluates the biased pdf formula q~ at
    # v.u' < 0 too, so the mean of the weights is the sphere integral of m HG/(p q~+(1-p)HG),
    # m the realised mixture: 1 on the hemisphere towards the probe, less on the other.
    p = spinbias_params()
    g, a = 0.9, 0.925
    u0 = np.array([-0.4, 0.1, 0.5])
    u0 /= np.linalg.norm(u0)
    pos = np.array([0.03, -0.02, 0.02])  # v is far from vertical: stays F = 2
    state = [*pos, *u0, 0.5, 1.0, 1, 0, g, 101.0]
    n = 400000
    out = OR.spin_bias_batch(p, state, (0.0, 0.0, 0.0), 11, table, n)
    assert np.all(out[:, 4] == 2)
    sel = out[:, 5]
    assert abs((sel == 1).mean() - 0.5) < 0.005
    f = out[:, 3]
    v = -pos / np.linalg.norm(pos)
    w = out[:, :3]
    denom = 0.5 * bias_pdf(w @ v, a) + 0.5 * hg_pdf(w @ u0, g)
    assert np.allclose(f, hg_pdf(w @ u0, g) / denom, rtol=1e-9)

    def weighted(W):
        mv, mu = W @ v, W @ u0
        h = hg_pdf(mu, g) / (2 * math.pi)
        q = np.where(mv >= 0, bias_pdf(np.clip(mv, 0, 1), a), 0.0) / (2 * math.pi)
        qt = bias_pdf(mv, a) / (2 * math.pi)
        m = 0.5 * q + 0.5 * h
        lr = h / (0.5 * qt + 0.5 * h)
        return np.column_stack([m * lr, m * lr * W[:, 0], m * lr * W[:, 1], m * lr * W[:, 2]])

    expect = sphere_integral(u0, weighted, g)
    assert 0.5 < expect[0] < 1.0
    assert abs(f.mean() - expect[0]) < 5 * f.std() / math.sqrt(n), (f.mean(), expect[0])
    mw = (f[:, None] * w).mean(axis=0)
    sw = (f[:, None] * w).std(axis=0) / math.sqrt(n)
    assert np.all(np.abs(mw - expect[1:]) < 5 * sw), (mw, expect[1:])


def test_plain_hg_spin_mean_direction(table):
    # TypeBias 0: Spin around u0 with HG; E[u'] = g u0 and |u'| = 1.
    p = spinbias_params()
    g = 0.8
    u0 = np.array([0.2, 0.6, -0.5])
    u0 /= np.linalg.norm(u0)
    state = [0.0, 0.0, 0.2, *u0, 0.5, 1.0, 0, 0, g, 101.0]  # z outside target depth: F = 0
    n = 200000
    out = OR.spin_bias_batch(p, state, (0.0, 0.0, 0.0), 3, table, n)
    assert np.all(out[:, 4] == 0) and np.all(out[:, 3] == 1.0)
    w = out[:, :3]
    assert np.allclose(np.linalg.norm(w, axis=1), 1, atol=1e-12)
    m, s = w.mean(axis=0), w.std(axis=0) / math.sqrt(n)
    assert np.all(np.abs(m - g * u0) < 5 * s)


# ------------------------------------------------------------------ HitBoundary
def tet_planes(V):
    faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    c = V.mean(axis=0)
    out = []
    for f in faces:
        a, b, d = V[list(f)]
        n = np.cross(b - a, d - a)
        n /= np.linalg.norm(n)
        off = -n @ a
        if n @ c + off < 0:
            n, off = -n, -off
        out.append([*n, off])
    return np.array(out)


def test_hit_boundary_against_barycentric_exit():
    rng = np.random.default_rng(2)
    for _ in range(500):
        V = rng.normal(size=(4, 3))
        if abs(np.linalg.det(V[1:] - V[0])) < 0.1:
            continue
        P = tet_planes(V)
        lam = rng.dirichlet(np.ones(4))
        x = lam @ V
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        # barycentric coordinates along the ray: lam(t) = lam + t dlam, exit at first zero
        M = np.vstack([V.T, np.ones(4)])
        dl = np.linalg.solve(M, np.append(u, 0.0))
        ts = [(-lam[i] / dl[i], i) for i in range(4) if dl[i] < 0]
        t_exit, vtx = min(ts)
        face_of_vertex = {3: 0, 2: 1, 1: 2, 0: 3}[vtx]  # face j is opposite vertex 3-j
        out = OR.hit_boundary(P.ravel(), x, u, 1e9, 2.0)
        assert out[0] == 1 and int(out[1]) == face_of_vertex
        assert abs(out[2] - t_exit) < 1e-9 * max(1, t_exit)
        assert abs(out[3] - (1e9 - out[2]) * 2.0) < 1e-3
        short = OR.hit_boundary(P.ravel(), x, u, 0.5 * t_exit, 2.0)
        assert short[0] == 0 and short[1] == -1 and short[2] == 0.5 * t_exit


# ------------------------------------------------------------------ whole path
def test_glass_slab_known_answer_oracle(table):
    d, h, cells = 0.1, 0.1 / 8, 8
    mesh = O.Mesh.generate(O.GEOM_HOMOGENEOUS, (cells, cells, cells), (cells * h, cells * h, d))
    params = O.make_params([(1.5, 0.0, 0.0, 0.9)], num_bins=900, depth=0.27)
    x0, y0 = 0.13 * h, 0.07 * h
    root, z = mesh.find_root(x0, y0, 0.0)
    n = 4000
    r = OR.run_ascan(mesh.soa(), params, O.ascan(x0, y0, z, root, n), 1, table, threads=4)
    R = ((1.5 - 1.0) / 2.5) ** 2
    c1 = (1 - R) ** 2 * R
    w = params.coherence_length_source / 6
    vp = int(np.floor(d / w))
    assert r["num_class1"] == n and r["killed"] == 0
    for b in range(vp - 3, vp + 3):
        assert r["count"][0][b] == n
        assert abs(r["sum"][0][b] / n - c1) < 1e-12
    assert r["count"][0][vp - 4] == 0 and r["count"][0][vp + 3] == 0
    vp2 = int(np.floor(2 * d / w))
    m2 = r["sum"][1][vp2] / n
    se = math.sqrt((r["sumsq"][1][vp2] / n - m2 * m2) / (n - 1))
    assert abs(m2 - (1 - R) ** 2 * R ** 3) < 5 * se


def test_type_bias_zero_tallies_every_aperture_hit(table):
    # TypeBias 0 has no filter: every upward boundary hit inside the aperture is detected.
    mesh = O.Mesh.generate(O.GEOM_HOMOGENEOUS, (8, 8, 8), (0.1, 0.1, 0.1))
    params = O.make_params([(1.4, 1.0, 100.0, 0.9)], num_bins=128, depth=0.1, type_bias=0,
                           max_collecting_radius=0.05, max_collecting_angle_deg=30.0)
    root, z = mesh.find_root(0.0013, 0.0007, 0.0)
    r = OR.run_ascan(mesh.soa(), params, O.ascan(0.0013, 0.0007, z, root, 2000), 1, table, threads=4)
    assert r["killed"] == 0 and r["num_filtered"] > 0
    assert r["num_class1"] + r["num_class2"] <= r["num_filtered"]


def _bin_params(bins=20, shift=0.0):
    p = O.make_params([(1.4, 1.0, 100.0, 0.9)], num_bins=bins, depth=bins * 2.0 ** -10)
    p.optical_depth_shift = shift
    return p


# (OP/2 in bin units, bins receiving the contribution, Class counter bumped)
# from the reference's window [VP-3, VP+3) clipped to [0, B) and its `ii == VP` counter
# (src/octmps_kernel.cuh:684-741), B = 20, bin width 2^-10 cm.
BIN_CASES = [
    (10.5, [7, 8, 9, 10, 11, 12], True),
    (10.0, [7, 8, 9, 10, 11, 12], True),      # exact edge: VP = 10, not 9
    (0.5, [0, 1, 2], True),                    # window clipped at 0
    (19.5, [16, 17, 18, 19], True),            # clipped at B
    (21.0, [18, 19], False),                   # VP = 21 outside: bins 18, 19 only, no class count
    (30.0, [], False),                         # nothing in range, still one filtered detection
]


@pytest.mark.parametrize("half_op,bins,counted", BIN_CASES)
def test_tally_window_and_counters(half_op, bins, counted):
    p = _bin_params()
    w, rF, lr = 0.5, 0.25, 0.75
    c = w * (1 - rF) * lr
    op = 2 * half_op * 2.0 ** -10
    maxz = 1.0  # deep: Class I
    sums, counts, ctr = OR.tally_detection(p, op, maxz, w, rF, lr)
    want = np.zeros(20)
    want[bins] = c
    assert np.array_equal(sums[0], want) and not sums[1].any()
    assert np.array_equal(counts[0], (want > 0).astype(np.uint64))
    assert list(ctr) == [1, 1 if counted else 0, 0]


def test_tally_class_rule_is_strict_and_negative_positions_saturate():
    p = _bin_params()
    lc = p.coherence_length_source  # 6 * 2^-10: every quantity below is exact in binary
    op = 21 * 2.0 ** -10            # OP/2 = 10.5 bins
    edge = op / 2 - lc / 2          # Class I needs MaxDepth > OP/2 - Lc/2
    _, counts, ctr = OR.tally_detection(p, op, edge, 1.0, 0.0, 1.0)
    assert counts[1].sum() == 6 and counts[0].sum() == 0 and list(ctr) == [1, 0, 1]
    _, counts, ctr = OR.tally_detection(p, op, np.nextafter(edge, 1.0), 1.0, 0.0, 1.0)
    assert counts[0].sum() == 6 and list(ctr) == [1, 1, 0]
    # OP/2 below the shift: the position saturates to bin 0 (window 0..2).
    q = _bin_params(shift=5 * 2.0 ** -10)
    _, counts, ctr = OR.tally_detection(q, 4 * 2.0 ** -10, 1.0, 1.0, 0.0, 1.0)
    assert list(np.nonzero(counts[0])[0]) == [0, 1, 2] and list(ctr) == [1, 1, 0]
