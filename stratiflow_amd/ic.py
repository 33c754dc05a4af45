"""Seeded initial conditions and background profiles (host-side, numpy).

The reference seeds its perturbations from ``std::random_device``
(Field.h:1029-1031, 1074-1076), so its runs cannot be repeated.  Here every
perturbation is a pure function of a 64-bit seed, so the GPU path and the
CPU oracle receive byte-identical inputs.

Perturbation (per field u1, u2 (3D), u3, b):
  f(j1, j2, z) = sum_{h=0..3} c[j1, j2, h] * exp(-(z/w)^2) * H_h(z/w)
over all active modes (j1 < N1/3, active j2), with c = a + i*b and a, b
uniform in [-1, 1) from PCG64(seed), H_h the physicists' Hermite
polynomials and w the Gaussian window width.  The j1 = 0 column is made
Hermitian so the field is real.  Each field is scaled so that the maximum
absolute value of its nodal (physical-space) values equals ``amplitude``.
Neumann fields are evaluated at VerticalPointsFractional, the Dirichlet
field u3 at VerticalPoints (Field.h:862-871).
"""
import numpy as np

THREE_D, TWO_D, TWO_AND_A_HALF = 0, 1, 2


def hermite(h, x):
    """Physicists' Hermite polynomial H_h(x) (Field.cpp:91-131 computes the same family)."""
    if h == 0:
        return np.ones_like(x)
    hm, hc = np.ones_like(x), 2 * x
    for m in range(1, h):
        hm, hc = hc, 2 * x * hc - 2 * m * hm
    return hc


def active_j2(N2, dim):
    """Active spanwise indices in the reference's order (Field.h:1129-1166)."""
    if dim == TWO_D or N2 == 1:
        return list(range(N2))
    h = N2 // 3
    return list(range(h)) + list(range(N2 - h + 1, N2))


def neumann_1d(values):
    """Nodal1D::SetValue ghost rule for a Neumann profile (Field.h:655-682)."""
    v = np.array(values, dtype=np.float64)
    v[1] = v[2]
    v[0] = v[1]
    v[-2] = v[-3]
    v[-1] = v[-2]
    return v


def tanh_background(zf):
    """U_ = tanh(z) set with Neumann1D::SetValue (IMEXRK.h:326-332, Parameters.h:33-36)."""
    return neumann_1d(np.tanh(zf))


def tanh_stratification(zf):
    """dB/dz of B = tanh(z) at the Neumann points."""
    return 1.0 / np.cosh(zf) ** 2


def nodal_of(modal, N1):
    """Unnormalised c2r of a (N2, M1, N3) spectrum, as ModalField::ToNodal (Field.h:965-994)."""
    N2 = modal.shape[0]
    return np.fft.irfft2(modal, s=(N2, N1), axes=(0, 1)) * (N1 * N2)


def perturbation(N1, N2, N3, dim, z, zf, seed, amplitude=1e-3, width=2.0, fields=None):
    """Return {name: modal array (N2, N1//2+1, N3) complex128} for u1, u2, u3, b."""
    M1 = N1 // 2 + 1
    A1 = N1 // 3
    js = active_j2(N2, dim)
    if fields is None:
        fields = ["u1", "u2", "u3", "b"] if dim != TWO_D else ["u1", "u3", "b"]
    rng = np.random.Generator(np.random.PCG64(seed))
    out = {}
    for name in fields:
        zz = z if name == "u3" else zf
        x = zz / width
        prof = np.stack([np.exp(-x ** 2) * hermite(h, x) for h in range(4)])  # (4, N3)
        m = np.zeros((N2, M1, N3), dtype=np.complex128)
        c = rng.uniform(-1.0, 1.0, size=(len(js), A1, 4)) + 1j * rng.uniform(-1.0, 1.0, size=(len(js), A1, 4))
        for a, j2 in enumerate(js):
            m[j2, :A1, :] = c[a] @ prof
        # Hermitian j1 = 0 column: f(0, -k) = conj f(0, k), f(0, 0) real
        if N2 > 1:
            for j2 in js:
                if 0 < j2 < N2 // 2 + 1:
                    m[(N2 - j2) % N2, 0, :] = np.conj(m[j2, 0, :])
        m[0, 0, :] = m[0, 0, :].real
        mx = np.abs(nodal_of(m, N1)).max()
        if mx > 0:
            m *= amplitude / mx
        out[name] = m
    return out
