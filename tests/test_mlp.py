"""Warm-start network: the CPU oracle (oracle/mlp_oracle.py) pinned by finite
differences, the GPU kernels (BF16 MFMA forward/backward, weight gradients)
against the oracle, the GPU features against a host computation, and the gain
of the trained network over the MBPT2 start."""
import numpy as np
import pytest

import mlp_oracle as mo


def random_problem(seed, n):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, np.sqrt(2 / 16), (16, 64)), "b1": rng.normal(0, 0.1, 64),
         "W2": rng.normal(0, np.sqrt(2 / 64), (64, 64)), "b2": rng.normal(0, 0.1, 64),
         "W3": rng.normal(0, np.sqrt(2 / 64), (64, 64)), "b3": rng.normal(0, 0.1, 64),
         "w4": rng.normal(0, 0.3, 64), "b4": rng.normal(0, 0.1, 1)}
    theta = mo.pack(p).astype(np.float32).astype(np.float64)
    x = rng.normal(0, 1, (n, 16)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    return theta, x, t


def test_oracle_gradient_by_finite_differences():
    theta, x, t = random_problem(1, 37)
    x = x.astype(np.float64)
    t = t.astype(np.float64)
    _, g = mo.loss_grad(theta, x, t)
    rng = np.random.default_rng(2)
    # a few parameters of every block (offsets of the flat layout)
    offs, o = [], 0
    for _, shp in mo.SHAPES:
        n = int(np.prod(shp))
        offs += list(o + rng.choice(n, min(n, 12), replace=False))
        o += n
    h = 1e-6
    for k in offs:
        e = np.zeros_like(theta)
        e[k] = h
        fd = (mo.loss(theta + e, x, t) - mo.loss(theta - e, x, t)) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_gelu_polynomial_cdf_against_exact_cdf():
    """The network's GELU takes the normal CDF as an odd polynomial on [-4, 4]
    (the kernels' definition, restated by the oracle): within 6e-5 of the
    exact CDF everywhere, far below the BF16 rounding of the activations
    (2^-9 relative), odd about 1/2, and gelu_grad its derivative."""
    from scipy.special import erf
    z = np.linspace(-12.0, 12.0, 240001)
    exact = 0.5 * (1.0 + erf(z / np.sqrt(2.0)))
    assert np.max(np.abs(mo.phi(z) - exact)) < 6e-5
    assert np.max(np.abs(mo.phi(z) + mo.phi(-z) - 1.0)) < 1e-15
    # a plain erf-free polynomial: no sign error or dropped term hides behind
    # the clamp (Phi(0) = 1/2, slope 1/sqrt(2 pi) at 0)
    assert mo.phi(0.0) == 0.5
    assert abs(mo.gelu_grad(0.0) - 0.5) < 1e-15
    h = 1e-6
    zz = np.concatenate([np.linspace(-3.99, 3.99, 801), [-6.0, 5.0]])
    fd = (mo.gelu(zz + h) - mo.gelu(zz - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(zz))) < 1e-7
    d0 = (mo.phi(h) - mo.phi(-h)) / (2 * h)
    assert abs(d0 - 1.0 / np.sqrt(2.0 * np.pi)) < 1e-4


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7
    vals = np.array([one, one + ulp / 2, one + 1.5 * ulp, one + ulp / 2 + 2 ** -20, -3.0 - ulp])
    want = np.array([one, one, one + 2 * ulp, one + ulp, -3.0])
    assert np.array_equal(mo.bf16(vals), want)


@pytest.mark.gpu
def test_mlp_forward_matches_oracle(ccnm):
    theta, x, t = random_problem(3, 1000)   # not a multiple of 64: padding path
    net = ccnm.WarmStart(seed=1)
    net.params = theta.astype(np.float32)
    y = net.forward(x)
    yo, _ = mo.forward(theta, x.astype(np.float64), bf=True)
    assert np.max(np.abs(y - yo)) < 2e-3 * max(1.0, np.max(np.abs(yo))), np.max(np.abs(y - yo))


@pytest.mark.gpu
def test_mlp_gradient_matches_oracle(ccnm):
    """GPU loss and gradient (BF16 MFMA, FP32 sums) against the oracle with the
    same BF16 rounding points (FP64 sums)."""
    theta, x, t = random_problem(4, 4000)
    net = ccnm.WarmStart(seed=1)
    net.params = theta.astype(np.float32)
    loss, g = net.loss_grad(x, t)
    lo, go = mo.loss_grad(theta, x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss - lo) < 1e-4 * lo
    o = 0
    for name, shp in mo.SHAPES:
        n = int(np.prod(shp))
        a, b = g[o:o + n], go[o:o + n]
        assert np.linalg.norm(a - b) < 1e-2 * np.linalg.norm(b) + 1e-7, (name, np.linalg.norm(a - b),
                                                                          np.linalg.norm(b))
        o += n


def host_features(s, T0, twist):
    """Features of every stored amplitude from the solver's orbitals, labels and
    Fock diagonal (the definition documented in mlp.hip)."""
    orb = s.orbitals()
    lab = s.t2_labels()
    a, b, i, j = lab[:, 0], lab[:, 1], lab[:, 2], lab[:, 3]
    box = s.dims["box"]
    nvec = np.stack([orb["nx"], orb["ny"], orb["nz"]], 1).astype(np.float64)
    k = (2 * 3.14159265358979 * nvec + np.asarray(twist)[None, :]) / box
    n = s.dims["t2_unique"]
    f = s.debug_fetch("fdiag", s.dims["nsp"]).real
    kf = s.dims["kf"]
    hbarc, pmass = 197.326968, float(np.float32(939.565))
    ef = hbarc ** 2 * kf ** 2 / (2 * pmass)
    p = 0.5 * (k[a] - k[b])
    q = 0.5 * (k[i] - k[j])
    K = k[a] + k[b]
    npn, nq = np.linalg.norm(p, axis=1), np.linalg.norm(q, axis=1)
    den = f[i] + f[j] - f[a] - f[b]
    out = np.zeros((n, 16))
    out[:, 0] = npn / kf
    out[:, 1] = nq / kf
    out[:, 2] = np.linalg.norm(K, axis=1) / kf
    with np.errstate(invalid="ignore", divide="ignore"):
        out[:, 3] = np.where((npn > 0) & (nq > 0), np.sum(p * q, 1) / (npn * nq), 0.0)
    out[:, 4] = np.linalg.norm(p - q, axis=1) / kf
    out[:, 5], out[:, 6], out[:, 7], out[:, 8] = orb["sz"][a], orb["sz"][b], orb["sz"][i], orb["sz"][j]
    out[:, 9] = 0.5 * (orb["tz"][a] + orb["tz"][b])
    out[:, 10] = kf
    out[:, 11], out[:, 12], out[:, 13] = np.asarray(twist) / 3.14159265358979
    out[:, 14] = 25 * T0.real
    out[:, 15] = den / ef
    return out


@pytest.mark.gpu
def test_mlp_features_match_host(ccnm):
    twist = (0.4, 1.3, 2.2)
    s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.12, nmax=2, twist=twist)
    s.generate_interaction()
    s.setup_t_amplitudes()
    T0 = s.debug_fetch("T", s.dims["t2_unique"])
    net = ccnm.WarmStart(seed=1)
    f, tgt = net.features(s)
    want = mo.bf16(host_features(s, T0, twist))
    # one BF16 step at most (the device computes the inputs in double, then
    # rounds once)
    tol = 2.0 ** -7 * np.abs(want) + 1e-6
    bad = np.abs(f - want) > tol
    cols = np.nonzero(bad.any(0))[0]
    assert not bad.any(), [(c, int(bad[:, c].sum()), f[bad[:, c], c][:3], want[bad[:, c], c][:3]) for c in cols]
    # targets of the MBPT2 amplitudes themselves are zero
    assert np.max(np.abs(tgt)) == 0.0


@pytest.mark.gpu
def test_warm_start_gain_and_energy_independence(ccnm):
    """Train on converged solves of |n_i|<=2 lattices (random densities and
    twists), then on held-out lattices (other densities and twists, and a
    larger |n_i|<=3 lattice): the learned start is much closer to the converged
    amplitudes than MBPT2, and the converged energy does not depend on the
    start (the FP64 fixed point decides it)."""
    rng = np.random.default_rng(5)
    net = ccnm.WarmStart(seed=3)
    for _ in range(8):
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=rng.uniform(0.06, 0.22), nmax=2,
                        twist=tuple(rng.uniform(0, np.pi, 3)))
        s.generate_interaction()
        s.setup_t_amplitudes()
        assert s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)["converged"]
        net.add_solve(s)
    first = net.train(50, batch=8192, lr=2e-3, seed=1)
    last = net.train(550, batch=8192, lr=1e-3, seed=2)
    assert last < 0.5 * first, (first, last)
    for nmax, rho, tw in ((2, 0.11, (0.5, 2.5, 1.5)), (2, 0.19, (2.8, 0.3, 1.0)), (3, 0.15, (1.2, 0.4, 2.0))):
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=rho, nmax=nmax, twist=tw)
        s.generate_interaction()
        n = s.dims["t2_unique"]
        s.setup_t_amplitudes()
        t_mbpt2 = s.debug_fetch("T", n)
        ref = s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)
        t_conv = s.debug_fetch("T", n)
        net.warm_start(s)
        t_learn = s.debug_fetch("T", n)
        out = s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)
        err_m = np.linalg.norm(t_mbpt2 - t_conv) / np.linalg.norm(t_conv)
        err_l = np.linalg.norm(t_learn - t_conv) / np.linalg.norm(t_conv)
        print(f"nmax {nmax} rho {rho}: start error MBPT2 {err_m:.3f} learned {err_l:.3f}, "
              f"iterations {ref['iterations']} -> {out['iterations']}")
        assert err_l < 0.6 * err_m
        assert out["converged"]
        # both stop at |dE| <= 1e-8 MeV per iteration; the fixed points agree to
        # well below the 1e-5 MeV per nucleon accuracy target
        assert abs(out["eccd"] - ref["eccd"]) / 66 < 1e-6


@pytest.mark.gpu
def test_warm_start_cuts_iterations_on_held_out_lattices(ccnm):
    """The warm-start protocol at reduced size: train on 24 converged solves of
    PNM A=66 |n_i|<=2 with the chiral NNLO(450) NN at random densities, twists
    and c1, c3 (+-20%), then solve 6 held-out |n_i|<=3 lattices drawn the same
    way (other seed) from MBPT2 and from the learned start with the reference's
    update (DIIS 8/8, FP64, |dE| <= 1e-8 MeV): the learned start needs fewer
    iterations on average and reaches the same energy."""
    rng = np.random.default_rng(2718)
    ci0 = (-0.74, -0.49, -0.65, 0.96)

    def lattice(nmax):
        ci = list(ci0)
        ci[0] *= rng.uniform(0.8, 1.2)
        ci[2] *= rng.uniform(0.8, 1.2)
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=rng.uniform(0.06, 0.22), nmax=nmax,
                        twist=tuple(rng.uniform(0, np.pi, 3)))
        s.set_chiral(ccnm.NNLO, 450.0, ci)
        s.generate_interaction(ccnm.INT_CHIRAL)
        return s

    net = ccnm.WarmStart(seed=21)
    for _ in range(24):
        s = lattice(2)
        s.setup_t_amplitudes()
        assert s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100, accel=1, mix=1.0)["converged"]
        net.add_solve(s)
        s.close()
    net.train(1500, batch=16384, lr=2e-3, seed=1)
    net.train(500, batch=16384, lr=5e-4, seed=2)
    it_m, it_l = [], []
    for _ in range(6):
        s = lattice(3)
        s.setup_t_amplitudes()
        rm = s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)
        net.warm_start(s)
        rl = s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)
        assert rm["converged"] and rl["converged"]
        assert abs(rm["eccd"].real - rl["eccd"].real) / 66 < 1e-6
        it_m.append(rm["iterations"])
        it_l.append(rl["iterations"])
        s.close()
    print("iterations MBPT2", it_m, "learned", it_l)
    assert np.mean(it_l) < np.mean(it_m), (it_m, it_l)


@pytest.mark.gpu
def test_fused_and_batched_warm_start_equal_the_three_kernel_path(ccnm, monkeypatch):
    """The fused start (features in registers, network, t = t_MBPT2 + scale y
    in one kernel) writes the same bits as the three kernels of the training
    path (features, network, apply; CCGPU_MLP_FUSED=0), and the batched entry
    point the same bits as one call per solve."""
    rng = np.random.default_rng(9)
    net = ccnm.WarmStart(seed=4)
    s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.13, nmax=2, twist=(0.4, 1.3, 2.2))
    s.generate_interaction()
    s.setup_t_amplitudes()
    assert s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)["converged"]
    net.add_solve(s)
    net.train(100, batch=8192, lr=2e-3, seed=1)
    sols = [ccnm.Solver(A=66, matter=ccnm.PNM, rho=rho, nmax=2, twist=tuple(rng.uniform(0, np.pi, 3)))
            for rho in (0.09, 0.17, 0.21)]
    for x in sols:
        x.generate_interaction()
    out = {}
    for mode in ("0", "1", "batch"):
        monkeypatch.setenv("CCGPU_MLP_FUSED", "0" if mode == "0" else "1")
        for x in sols:
            x.setup_t_amplitudes()   # every start overwrites the MBPT2 amplitudes
        if mode == "batch":
            net.warm_start_batch(sols)
        else:
            for x in sols:
                net.warm_start(x)
        out[mode] = [x.debug_fetch("T", x.dims["t2_unique"]) for x in sols]
    for a, b, c in zip(out["0"], out["1"], out["batch"]):
        assert np.array_equal(a, b)
        assert np.array_equal(a, c)
        assert np.any(a != 0)
