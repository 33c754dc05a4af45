"""CPU restatement of the Krylov drivers over the oracle (test infrastructure only).

Follows NewtonKrylov.h:119-235 (finite-difference directional derivative,
modified Gram-Schmidt Arnoldi on A = -dF/dx, least squares in the singular
basis of the Hessenberg matrix with the trust region |z| <= Delta) and
Arnoldi.h:15-193 (Krylov basis of the linearised flow map, eigenvalues of the
Hessenberg matrix).  Vectors are dicts {name: modal array} in the oracle's
(N2, M1, N3) layout; every StateVector operation re-applies EnforceBCs
(StateVector.h:415-421) as the reference's do.  The dense parts use numpy's
LAPACK SVD and eigenvalue routines in FP64; they pin the GPU build's own
Jacobi SVD and Hessenberg QR, which are written from the same definitions.
"""
import numpy as np

NEUMANN, DIRICHLET = 0, 1


def names(three):
    return ["u1", "u2", "u3", "b"] if three else ["u1", "u3", "b"]


def enforce_bcs(x):
    """u3.ZeroEnds (Dirichlet: slots 0,1,2,N-2,N-1); u1,u2,b NeumannEnds (Field.h:475-500)."""
    for n, a in x.items():
        if n == "u3":
            a[..., [0, 1, 2, -2, -1]] = 0
        else:
            a[..., 0] = a[..., 1]
            a[..., -1] = a[..., -2]
    return x


def copy(x):
    return {n: a.copy() for n, a in x.items()}


def muladd(x, a, y):
    """x += a*y, then EnforceBCs (StateVector.h:67-78)."""
    for n in x:
        x[n] += a * y[n]
    return enforce_bcs(x)


def scale(x, a):
    for n in x:
        x[n] *= a
    return enforce_bcs(x)


def dot(o, x, y, Ri):
    """StateVector::Dot (StateVector.h:122-134) with InnerProd (Integration.h:117-146), FP64."""
    s = 0.0
    for n in x:
        w = Ri if n == "b" else 1.0
        s += w * o.inner_prod(x[n], y[n], DIRICHLET if n == "u3" else NEUMANN)
    return s


def norm(o, x, Ri):
    return np.sqrt(dot(o, x, x, Ri))


def _load_state(o, x):
    for n, a in x.items():
        o.set_modal(n, a)
    o.set_modal("p", np.zeros(o.modal_shape, complex))


def phi_fixed(o, x, dt, nsteps):
    """nsteps IMEX-RK steps at fixed dt after FixedEvolve's start-up (StateVector.cpp:85-105,
    with PrepareRun's populate, IMEXRK.h:164-168), p = 0."""
    _load_state(o, x)
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    o.set_dt(dt)
    o.populate_nodal()
    o.step(nsteps)
    return enforce_bcs({n: o.get_modal(n) for n in x})


def linear_evolve(o, x, about, T, real=np.float64):
    """StateVector::LinearEvolve (StateVector.cpp:107-156), p = 0: dt starts at 0.01,
    CFLlinear (IMEXRK.h:376-393) rescales it to CFL 0.4 every 100 steps, and when
    fewer than 100 steps remain it is shortened so the last step ends at T.  `real`
    is the driver's stratifloat, in which t, dt and the step count are kept."""
    tot = {"u1": "u1t", "u2": "u2t", "u3": "u3t", "b": "bt"}
    for n, a in about.items():
        o.set_modal(tot[n], a)
    o.set_linear_background()
    _load_state(o, x)
    o.filter_all()
    o.populate_nodal()
    o.remove_divergence(0.0)
    o.populate_nodal()  # PrepareRunLinear (IMEXRK.h:200-204)
    T = real(T)
    dt = real(0.01)
    o.set_dt(float(dt))
    t, step = real(0), 0
    while t + real(0.0001) < T:
        if step % 100 == 0:
            dt = real(dt * real(0.4) / real(o.cfl(linear=True)))
            o.set_dt(float(dt))
            remaining = real(T - t)
            steps_left = int(remaining / dt) + 1
            if steps_left < 100:
                dt = real(remaining / real(steps_left))
                o.set_dt(float(dt))
        o.step_linear(1)
        t = real(t + dt)
        step += 1
    return enforce_bcs({k: o.get_modal(k) for k in x})


def hookstep(H, beta, Delta):
    """Least squares min |H y - beta e1| in the singular basis of H, trust region |z| <= Delta (NewtonKrylov.h:179-218)."""
    U, d, Vt = np.linalg.svd(H, full_matrices=False)
    p = U[0, :] * beta
    with np.errstate(divide="ignore"):
        z = np.where(d > 0, p / d, np.inf)
    mu = 0.0
    if Delta > 0:
        while np.linalg.norm(z) > Delta:
            mu += 0.00001
            z = p * d / (d * d + mu)
    y = Vt.T @ z
    r = H @ y
    r[0] -= beta
    return y, np.linalg.norm(r) / beta, mu


def gmres(o, F, x, Fx, Ri, K, epsilon, Delta, fd_epsilon=1e-7):
    """One GMRES solve of (-dF/dx) dx = F(x) about x (NewtonKrylov.h:138-235).

    F maps a state dict to a state dict; Fx = F(x).  Returns dx, H and the
    GMRES residual history."""
    def deriv(v):
        eps = fd_epsilon * norm(o, x, Ri) / norm(o, v, Ri)
        t = muladd(copy(x), eps, v)
        t = F(t)
        t = muladd(t, -1.0, Fx)
        return scale(t, 1.0 / eps)

    q = [enforce_bcs(copy(Fx))]
    beta = norm(o, q[0], Ri)
    scale(q[0], 1.0 / beta)
    H = np.zeros((K, K - 1))
    hist = []
    Kused = K
    y = None
    for k in range(1, K):
        qk = scale(deriv(q[k - 1]), -1.0)
        for j in range(k):
            H[j, k - 1] = dot(o, q[j], qk, Ri)
            muladd(qk, -H[j, k - 1], q[j])
        H[k, k - 1] = norm(o, qk, Ri)
        scale(qk, 1.0 / H[k, k - 1])
        enforce_bcs(qk)
        q.append(qk)
        y, res, _ = hookstep(H[:k + 1, :k], beta, Delta)
        hist.append(res)
        if res < epsilon:
            Kused = k + 1
            break
    dx = {n: np.zeros_like(a) for n, a in x.items()}
    for k in range(Kused - 1):
        muladd(dx, y[k], q[k])
    return dx, H, hist


def arnoldi(o, L, q0, Ri, K):
    """Krylov basis of the linear map L and the eigenvalues of its Hessenberg matrix (Arnoldi.h:15-117)."""
    q = [enforce_bcs(copy(q0))]
    scale(q[0], 1.0 / norm(o, q[0], Ri))
    H = np.zeros((K, K - 1))
    for k in range(1, K):
        qk = L(q[k - 1])
        for j in range(k):
            H[j, k - 1] = dot(o, q[j], qk, Ri)
            muladd(qk, -H[j, k - 1], q[j])
        H[k, k - 1] = norm(o, qk, Ri)
        scale(qk, 1.0 / H[k, k - 1])
        enforce_bcs(qk)
        q.append(qk)
    ev = np.linalg.eigvals(H[:K - 1, :K - 1])
    order = np.lexsort((-ev.imag, -np.abs(ev)))
    return ev[order], H, q
