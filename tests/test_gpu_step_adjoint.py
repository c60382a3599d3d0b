"""Reverse-mode derivatives of the barotropic step (Barotropic_Dynamics.jl:2-46)
and the shallow-water step (Shallow_Water_Dynamics.jl:7-179).

Each linear operator's adjoint is pinned by the dot-product identity
<A x, y> = <x, A^T y> against the forward entry point on the device (FP64,
1e-12 relative).  The composition is pinned by central differences of a run
of the steppers themselves (the exp/Barotropic and exp/Shallow_Water loops,
Update_Init_Step! after the first step): along random directions the difference quotient of
J = sum of Re<W_f, f^K> over the final state fields must converge to the gradient at second
order in the step (a wrong gradient leaves a constant error) and agree with
it to 1e-6 at the smallest step.  The checkpoint interval must not change
the gradient at all (bit-exact)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

A, OMEGA = 6.371e6, 7.292e-5


def rspec(N, nlev, rng, upper_only=True, valid=False):
    s = rng.standard_normal((nlev, N + 2, N + 1)) + 1j * rng.standard_normal((nlev, N + 2, N + 1))
    if upper_only or valid:
        for m in range(N + 1):
            s[:, :m, m] = 0
    if valid:  # a spectral state of a real field: m = 0 real, degree N+1 zero
        s[:, :, 0] = s[:, :, 0].real
        s[:, N + 1, :] = 0
    return s


def dot(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    if np.iscomplexobj(a):
        return float(np.sum(a.real * b.real + a.imag * b.imag))
    return float(np.sum(a * b))


def close(lhs, rhs, tol=1e-12):
    return abs(lhs - rhs) <= tol * (abs(lhs) + abs(rhs))


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), device="cuda")


@pytest.mark.parametrize("N,nlat", [(21, 32), (42, 64), (85, 128)])
def test_operator_dot_products(jgcm, N, nlat):
    from jgcm_amd import adjoint as ad
    nlev = 2
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nlev, A)
    rng = np.random.default_rng(N + 11)
    gu, gv = dev(rng.standard_normal((nlev, nlat, 2 * nlat))), dev(rng.standard_normal((nlev, nlat, 2 * nlat)))
    sa, sb = dev(rspec(N, nlev, rng)), dev(rspec(N, nlev, rng))
    # Vor_Div_From_Grid_UV!: (u, v) -> (vor, div)
    vor, div = mesh.spectral_zeros(), mesh.spectral_zeros()
    jgcm.Vor_Div_From_Grid_UV(mesh, gu, gv, vor, div)
    ub, vb = mesh.grid_zeros(), mesh.grid_zeros()
    ad.vor_div_from_grid_uv_adjoint(mesh, sa, sb, ub, vb)
    torch.cuda.synchronize()
    lhs, rhs = dot(vor, sa) + dot(div, sb), dot(gu, ub) + dot(gv, vb)
    assert close(lhs, rhs), (lhs, rhs)
    # vorticity part alone (div_bar = NULL)
    ad.vor_div_from_grid_uv_adjoint(mesh, sa, None, ub, vb)
    torch.cuda.synchronize()
    lhs, rhs = dot(vor, sa), dot(gu, ub) + dot(gv, vb)
    assert close(lhs, rhs), (lhs, rhs)
    # UV_Grid_From_Vor_Div!: (vor, div) -> (u, v)
    u, v = mesh.grid_zeros(), mesh.grid_zeros()
    jgcm.UV_Grid_From_Vor_Div(mesh, sa, sb, u, v)
    vb2, db2 = mesh.spectral_zeros(), mesh.spectral_zeros()
    ad.uv_grid_from_vor_div_adjoint(mesh, gu, gv, vb2, db2)
    torch.cuda.synchronize()
    lhs, rhs = dot(u, gu) + dot(v, gv), dot(sa, vb2) + dot(sb, db2)
    assert close(lhs, rhs), (lhs, rhs)
    # the u cos / v cos stencil over the whole array (lower triangle included)
    fa, fb = dev(rspec(N, nlev, rng, False)), dev(rspec(N, nlev, rng, False))
    ya, yb = dev(rspec(N, nlev, rng, False)), dev(rspec(N, nlev, rng, False))
    uc, vc = mesh.spectral_zeros(), mesh.spectral_zeros()
    jgcm.Compute_Ucos_Vcos_From_Vor_Div(mesh, fa, fb, uc, vc)
    xa, xb = mesh.spectral_zeros(), mesh.spectral_zeros()
    ad.ucos_vcos_adjoint(mesh, ya, yb, xa, xb)
    torch.cuda.synchronize()
    lhs, rhs = dot(uc, ya) + dot(vc, yb), dot(fa, xa) + dot(fb, xb)
    assert close(lhs, rhs), (lhs, rhs)


def test_vorticity_flux_adjoint_is_transposed_linearisation(jgcm):
    """The flux is quadratic in (ζ, u, v), so the central difference of the
    forward kernel is its exact linearisation (up to rounding)."""
    from jgcm_amd import adjoint as ad
    from jgcm_amd._lib import call
    from jgcm_amd.mesh import _ptr
    N, nlat = 42, 64
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    rng = np.random.default_rng(5)
    g = lambda: dev(rng.standard_normal((1, nlat, 2 * nlat)))
    z, u, v, dz, du, dv, yu, yv = (g() for _ in range(8))

    def flux(zz, uu, vv):
        av, fu, fv = mesh.grid_zeros(1), mesh.grid_zeros(1), mesh.grid_zeros(1)
        call("jgcm_vorticity_flux", mesh.handle, _ptr(zz), _ptr(uu), _ptr(vv), OMEGA, _ptr(av),
             _ptr(fu), _ptr(fv), 1, 0)
        return fu, fv

    h = 0.5
    fu1, fv1 = flux(z + h * dz, u + h * du, v + h * dv)
    fu0, fv0 = flux(z - h * dz, u - h * du, v - h * dv)
    ju, jv = (fu1 - fu0) / (2 * h), (fv1 - fv0) / (2 * h)
    zb, ub, vb = mesh.grid_zeros(1), mesh.grid_zeros(1), mesh.grid_zeros(1)
    ad.vorticity_flux_adjoint(mesh, z, u, v, OMEGA, yu, yv, zb, ub, vb)
    torch.cuda.synchronize()
    lhs, rhs = dot(ju, yu) + dot(jv, yv), dot(dz, zb) + dot(du, ub) + dot(dv, vb)
    assert close(lhs, rhs, 1e-12), (lhs, rhs)


@pytest.mark.parametrize("init", [True, False])
def test_damped_leapfrog_adjoint(jgcm, init):
    from jgcm_amd import adjoint as ad
    N, nlat = 42, 64
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True, 0,
                               864000)
    if not init:
        jgcm.Update_Init_Step(I)
    rng = np.random.default_rng(17)
    p, c, G = (dev(rspec(N, 1, rng, False)) for _ in range(3))
    P, C = dev(rspec(N, 1, rng, False)), dev(rspec(N, 1, rng, False))
    # forward: damping + leapfrog; then Time_Advance gives p' = filtered c, c' = n
    q, qp, qc, qn = G.clone(), p.clone(), c.clone(), mesh.spectral_zeros(1)
    jgcm.Damped_Leapfrog(I, q, qp, qc, qn)
    pb, cb, gb = mesh.spectral_zeros(1), mesh.spectral_zeros(1), mesh.spectral_zeros(1)
    ad.damped_leapfrog_adjoint(I, init, P, C, pb, cb, gb)
    torch.cuda.synchronize()
    lhs = dot(qc, P) + dot(qn, C)
    rhs = dot(p, pb) + dot(c, cb) + dot(G, gb)
    assert close(lhs, rhs), (lhs, rhs)
    if init:
        assert not pb.any()
    # in place (p_bar == p_bar_next, c_bar == c_bar_next) gives the same values
    P2, C2, g2 = P.clone(), C.clone(), mesh.spectral_zeros(1)
    ad.damped_leapfrog_adjoint(I, init, P2, C2, P2, C2, g2)
    torch.cuda.synchronize()
    assert torch.equal(P2, pb) and torch.equal(C2, cb) and torch.equal(g2, gb)


@pytest.mark.parametrize("init", [True, False])
def test_damped_leapfrog_adjoint_batch_equals_single_calls(jgcm, init):
    """One launch for several fields (levels differ, one field undamped, in
    place) gives the single calls' results bit for bit."""
    from jgcm_amd import adjoint as ad
    N, nlat = 42, 64
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 3, A)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True, 0,
                               864000)
    if not init:
        jgcm.Update_Init_Step(I)
    rng = np.random.default_rng(23)
    levels, damps = (3, 1, 2), (True, False, True)
    P = [dev(rspec(N, n, rng, False)) for n in levels]
    C = [dev(rspec(N, n, rng, False)) for n in levels]
    single = []
    for p, c, d in zip(P, C, damps):
        p2, c2, g2 = p.clone(), c.clone(), mesh.spectral_zeros(p.shape[0])
        ad.damped_leapfrog_adjoint(I, init, p2, c2, p2, c2, g2, damp=d)
        single.append((p2, c2, g2))
    G = [mesh.spectral_zeros(n) for n in levels]
    ad.damped_leapfrog_adjoint_batch(I, init, [(p, c, p, c, g, d)
                                               for p, c, g, d in zip(P, C, G, damps)])
    torch.cuda.synchronize()
    for (p2, c2, g2), p, c, g in zip(single, P, C, G):
        assert torch.equal(p, p2) and torch.equal(c, c2) and torch.equal(g, g2)


def rh4_state(jgcm, mesh):
    """Rossby-Haurwitz wave 4 vorticity (exp/Barotropic), spectral."""
    w = K = 7.848e-6
    mu = np.asarray(mesh.sinθ)[:, None]
    c = np.asarray(mesh.cosθ)[:, None]
    lam = np.asarray(mesh.λc)[None, :]
    z = (2 * w * mu - K * mu * c ** 4 * 30 * np.cos(4 * lam))[None]
    s = mesh.spectral_zeros(1)
    jgcm.Trans_Grid_To_Spherical(mesh, dev(z), s)
    return s


def run_model(jgcm, mesh, atmo, make_integrator, vor_p, vor_c, nsteps):
    """The exp/Barotropic time loop on the library's stepper."""
    N = mesh.num_fourier
    d = jgcm.Dyn_Data("b", N, N + 1, mesh.nλ, mesh.nθ, 1, mesh=mesh)
    d.spe_vor_p.copy_(vor_p)
    d.spe_vor_c.copy_(vor_c)
    jgcm.Trans_Spherical_To_Grid(mesh, d.spe_vor_c, d.grid_vor)
    jgcm.UV_Grid_From_Vor_Div(mesh, d.spe_vor_c, d.spe_zeros, d.grid_u_c, d.grid_v_c)
    I = make_integrator()
    for k in range(nsteps):
        jgcm.Barotropic_Dynamics(mesh, atmo, d, I)
        if I.init_step:
            jgcm.Update_Init_Step(I)
    torch.cuda.synchronize()
    return d.spe_vor_p.clone(), d.spe_vor_c.clone()


@pytest.mark.parametrize("starts_init", [True, False])
def test_barotropic_gradient_matches_central_differences(jgcm, starts_init):
    N, nlat, K = 42, 64, 12
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)

    def make_integrator():
        I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True,
                                   0, 864000)
        if not starts_init:
            jgcm.Update_Init_Step(I)
        return I

    rng = np.random.default_rng(23)
    c0 = rh4_state(jgcm, mesh)
    scale = float(c0.abs().max())
    p0 = c0 + dev(1e-2 * scale * rspec(N, 1, rng, valid=True))
    W = dev(rspec(N, 1, rng, valid=True))
    V = dev(rspec(N, 1, rng, valid=True))

    def J(vp, vc):
        pK, cK = run_model(jgcm, mesh, atmo, make_integrator, vp, vc, K)
        return dot(W, cK) + dot(V, pK), (pK, cK)

    adj = jgcm.Barotropic_Adjoint(mesh, atmo, make_integrator(), checkpoint_every=5)
    pK, cK = adj.forward(c0, p0, K)
    J0, (pK_ref, cK_ref) = J(p0, c0)
    assert torch.equal(pK, pK_ref) and torch.equal(cK, cK_ref)
    gp, gc = adj.backward(V, W)
    if starts_init:
        assert not gp.any()   # the first step never reads vor_p
    for seed in range(3):
        r = np.random.default_rng(100 + seed)
        dc = dev(rspec(N, 1, r, valid=True))
        dp = dev(rspec(N, 1, r, valid=True)) if not starts_init else torch.zeros_like(dc)
        dc, dp = dc / dc.abs().max(), dp / max(float(dp.abs().max()), 1.0)
        an = dot(gc, dc) + dot(gp, dp)
        err = []
        for f in (1e-3, 1e-4, 1e-5):
            h = f * scale
            Jp, _ = J(p0 + h * dp, c0 + h * dc)
            Jm, _ = J(p0 - h * dp, c0 - h * dc)
            err.append((Jp - Jm) / (2 * h) - an)
        tol = max(1.0, abs(an))
        # the quotient converges to the gradient at second order (a wrong
        # gradient leaves an error that does not shrink with h) ...
        assert abs(err[1]) <= 0.05 * abs(err[0]) + 1e-8 * tol, (seed, an, err)
        # ... and agrees with it closely at the smallest step
        assert abs(err[2]) <= 1e-6 * tol, (seed, an, err)


def test_checkpoint_interval_does_not_change_gradient(jgcm):
    N, nlat, K = 21, 32, 9
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True, 0,
                               864000)
    rng = np.random.default_rng(31)
    c0 = rh4_state(jgcm, mesh)
    W = dev(rspec(N, 1, rng, valid=True))
    zero = torch.zeros_like(W)
    grads = []
    for every in (1, 4, K, 100):
        adj = jgcm.Barotropic_Adjoint(mesh, atmo, I, checkpoint_every=every)
        adj.forward(c0, c0, K)
        grads.append(adj.backward(zero, W))
    assert I.init_step   # the integrator passed in is not advanced
    for gp, gc in grads[1:]:
        assert torch.equal(gp, grads[0][0]) and torch.equal(gc, grads[0][1])


def sw_case(jgcm, N, nlat):
    """A shallow-water state with flow, divergence and height structure
    (RH4 vorticity, a small random divergence, h from the exp/Shallow_Water
    forcing profile) and that forcing (SW.jl:57-77)."""
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    atmo = jgcm.Atmo_Data("sw", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)
    th = np.asarray(mesh.θc)[:, None]
    lam = np.asarray(mesh.λc)[None, :]
    h_0 = 3.0e4
    h_amp, h_lon, h_lat, h_width, h_itcz, itcz_w = (2.0e4, np.pi / 2, 25 * np.pi / 180,
                                                    15 * np.pi / 180, 1.0e5, 4 * np.pi / 180)
    x2 = ((lam - h_lon) / (2.0 * h_width)) ** 2
    y2 = ((th - h_lat) / h_width) ** 2
    h_eq = dev((h_0 + h_amp * np.exp(-(x2 + y2)) + h_itcz * np.exp(-(th / itcz_w) ** 2))[None])
    rng = np.random.default_rng(41)
    vor = rh4_state(jgcm, mesh)
    div = dev(1e-2 * float(vor.abs().max()) * rspec(N, 1, rng, valid=True))
    h = mesh.spectral_zeros(1)
    jgcm.Trans_Grid_To_Spherical(mesh, h_eq, h)
    state_c = [vor, div, h]
    state_p = [x + dev(1e-3 * float(x.abs().max()) * rspec(N, 1, rng, valid=True)) for x in state_c]
    state = (state_p[0], state_c[0], state_p[1], state_c[1], state_p[2], state_c[2])
    return mesh, atmo, h_0, h_eq, state


def run_sw_model(jgcm, mesh, atmo, make_integrator, h_0, h_eq, state, nsteps):
    """The exp/Shallow_Water time loop on the library's stepper, from a
    spectral state with its grid fields rebuilt by the inverse transforms."""
    N = mesh.num_fourier
    km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)
    d = jgcm.Dyn_Data("sw", N, N + 1, mesh.nλ, mesh.nθ, 1, mesh=mesh)
    names = ("spe_vor_p", "spe_vor_c", "spe_div_p", "spe_div_c", "spe_lnps_p", "spe_lnps_c")
    for n, t in zip(names, state):
        getattr(d, n).copy_(t)
    d.spe_lnps_n.copy_(d.spe_lnps_c)
    jgcm.Trans_Spherical_To_Grid(mesh, d.spe_vor_c, d.grid_vor)
    jgcm.Trans_Spherical_To_Grid(mesh, d.spe_div_c, d.grid_div)
    jgcm.Trans_Spherical_To_Grid(mesh, d.spe_lnps_c, d.grid_ps_c)
    jgcm.UV_Grid_From_Vor_Div(mesh, d.spe_vor_c, d.spe_div_c, d.grid_u_c, d.grid_v_c)
    I = make_integrator()
    for k in range(nsteps):
        jgcm.Shallow_Water_Physics(d, km, kt, h_eq)
        jgcm.Shallow_Water_Dynamics(mesh, atmo, h_0, d, I)
        if I.init_step:
            jgcm.Update_Init_Step(I)
    torch.cuda.synchronize()
    return tuple(getattr(d, n).clone() for n in names)


@pytest.mark.parametrize("K,starts_init", [(1, True), (1, False), (7, True)])
def test_shallow_water_gradient_matches_central_differences(jgcm, K, starts_init):
    N, nlat = 42, 64
    mesh, atmo, h_0, h_eq, state = sw_case(jgcm, N, nlat)
    km, kt = 1 / (20 * 86400.0), 1 / (10 * 86400.0)

    def make_integrator():
        I = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, mesh.laplacian_eig, 0.5, 1200, True, 0, 691200)
        if not starts_init:
            jgcm.Update_Init_Step(I)
        return I

    rng = np.random.default_rng(43)
    scales = [float(t.abs().max()) for t in state]
    weights = [dev(rspec(N, 1, rng, valid=True)) / s for s in scales]

    def J(st):
        out = run_sw_model(jgcm, mesh, atmo, make_integrator, h_0, h_eq, st, K)
        return sum(dot(w, o) for w, o in zip(weights, out)), out

    adj = jgcm.Shallow_Water_Adjoint(mesh, atmo, make_integrator(), h_0, km, kt, h_eq,
                                     checkpoint_every=3)
    final = adj.forward(state, K)
    _, ref = J(state)
    for a, b in zip(final, ref):
        assert torch.equal(a, b)
    grads = adj.backward(weights)
    if starts_init:
        # the first step reads vor_p nowhere (div_p and h_p do enter it, through
        # the implicit correction, Shallow_Water_Dynamics.jl:43-44)
        assert not grads[0].any()
    for seed in range(3):
        r = np.random.default_rng(200 + seed)
        dirs = []
        for k, s in enumerate(scales):
            dk = dev(rspec(N, 1, r, valid=True))
            dirs.append(s * dk / dk.abs().max())
        an = sum(dot(g, dk) for g, dk in zip(grads, dirs))
        err = []
        for f in (1e-3, 1e-4, 1e-5):
            Jp, _ = J(tuple(x + f * dk for x, dk in zip(state, dirs)))
            Jm, _ = J(tuple(x - f * dk for x, dk in zip(state, dirs)))
            err.append((Jp - Jm) / (2 * f) - an)
        tol = max(1.0, abs(an))
        assert abs(err[1]) <= 0.05 * abs(err[0]) + 1e-8 * tol, (seed, an, err)
        assert abs(err[2]) <= 1e-6 * tol, (seed, an, err)


def test_run_autograd_function_equals_reverse_sweep(jgcm):
    """Run.apply (the rrule of the time loop) gives exactly the reverse sweep's gradient."""
    from jgcm_amd.adjoint import Run
    N, nlat, K = 21, 32, 5
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True, 0,
                               864000)
    rng = np.random.default_rng(71)
    c0 = rh4_state(jgcm, mesh)
    W = dev(rspec(N, 1, rng, valid=True))
    vp = c0.clone().requires_grad_(True)
    vc = c0.clone().requires_grad_(True)
    adj = jgcm.Barotropic_Adjoint(mesh, atmo, I, checkpoint_every=2)
    pK, cK = Run.apply(adj, K, vp, vc)
    loss = (W.conj() * cK).real.sum()
    loss.backward()
    ref = jgcm.Barotropic_Adjoint(mesh, atmo, I, checkpoint_every=2)
    ref.forward(c0, c0, K)
    gp, gc = ref.backward(torch.zeros_like(W), W)
    assert torch.equal(vc.grad, gc) and torch.equal(vp.grad, gp)


def test_two_runs_on_one_adjoint_keep_their_own_trajectories(jgcm):
    """Two Run.apply calls on the same adjoint object, then backward of the
    first: each apply's gradient is that of its own trajectory (the adjoint
    object keeps only its latest run; Run keeps each apply's checkpoints)."""
    from jgcm_amd.adjoint import Run
    N, nlat = 21, 32
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True, 0,
                               864000)
    rng = np.random.default_rng(73)
    c0 = rh4_state(jgcm, mesh)
    c1 = c0 + 1e-6 * dev(rspec(N, 1, rng, valid=True))
    W = dev(rspec(N, 1, rng, valid=True))
    adj = jgcm.Barotropic_Adjoint(mesh, atmo, I, checkpoint_every=2)
    runs = []
    for c, K in ((c0, 5), (c1, 3)):
        vp = c.clone().requires_grad_(True)
        vc = c.clone().requires_grad_(True)
        pK, cK = Run.apply(adj, K, vp, vc)
        runs.append((c, K, vp, vc, (W.conj() * cK).real.sum()))
    for c, K, vp, vc, loss in runs:          # the first run's backward comes after both applies
        loss.backward()
        ref = jgcm.Barotropic_Adjoint(mesh, atmo, I, checkpoint_every=2)
        ref.forward(c, c, K)
        gp, gc = ref.backward(torch.zeros_like(W), W)
        assert torch.equal(vc.grad, gc) and torch.equal(vp.grad, gp), K


@pytest.mark.parametrize("prec", ["split32", "bf16x3"])
def test_gradient_with_reduced_precision_legendre(jgcm, prec):
    """The reverse sweep on a split-FP32 / BF16x3 Legendre mesh stays within
    1e-6 (relative, max norm) of the FP64 gradient."""
    N, nlat, K = 42, 64, 8
    grads = {}
    for p in ("fp64", prec):
        mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A, precision=p)
        atmo = jgcm.Atmo_Data("b", 2 * nlat, nlat, 1, False, False, False, False, mesh.sinθ, A, OMEGA)
        I = jgcm.Filtered_Leapfrog(0.04, 4, 1.0 / (0.1 * 86400), mesh.laplacian_eig, 0.0, 1800, True,
                                   0, 864000)
        c0 = rh4_state(jgcm, mesh)
        W = dev(rspec(N, 1, np.random.default_rng(5), valid=True))
        adj = jgcm.Barotropic_Adjoint(mesh, atmo, I, checkpoint_every=4)
        adj.forward(c0, c0, K)
        grads[p] = adj.backward(torch.zeros_like(W), W)[1].cpu().numpy()
    g0, g1 = grads["fp64"], grads[prec]
    err = np.abs(g1 - g0).max() / np.abs(g0).max()
    assert err < 1e-6, err
