"""Synthetic HDG workloads (manufactured solutions on structured simplex meshes).

Each workload returns the inputs the reference's drivers build by hand in
tests/regression/HDG/*.cpp: the mesh, the model kind, the field map
(Tau, DiffusionTensor, Velocity, Dirichlet) and the source at the element
integration points.
"""
import math

import numpy as np

from .hdg import Mesh

PI = math.pi


def face_normals(m):
    """Unit normal of every (straight) face, outward from its first cell."""
    a = m.arrays()
    x = a['coords']
    faces, f2c, cells = a['faces'], a['face2cell'], a['cells']
    dim = x.shape[1]
    fv = x[faces[:, :dim]]                       # face vertices
    if dim == 2:
        t = fv[:, 1] - fv[:, 0]
        n = np.stack([t[:, 1], -t[:, 0]], axis=1)
    else:
        n = np.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    cv = x[cells[f2c[:, 0], :dim + 1]]           # first cell's vertices
    centroid = cv.mean(axis=1)
    s = np.sign(np.einsum('fd,fd->f', fv[:, 0] - centroid, n))
    return n * s[:, None]


def dirichlet_from(m, u_exact):
    a = m.arrays()
    g = np.zeros(a['faces'].shape)
    xb = a['coords'][a['faces'][a['boundary']]]
    g[a['boundary']] = u_exact(xb)
    return g


def u_sin2d(x):
    return np.sin(PI * x[..., 0]) * np.sin(PI * x[..., 1])


def f_sin2d(x, D=1.0):
    return 2.0 * PI * PI * D * u_sin2d(x)


def u_c3(x):
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    return np.exp(X + Y + Z) * np.sin(PI * X) * np.sin(PI * Y) * np.sin(PI * Z)


def grad_lap_c3(x):
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    e = np.exp(X + Y + Z)
    s = [np.sin(PI * X), np.sin(PI * Y), np.sin(PI * Z)]
    c = [np.cos(PI * X), np.cos(PI * Y), np.cos(PI * Z)]
    g = []
    lap = 0.0
    for k in range(3):
        o = [s[j] for j in range(3) if j != k]
        # d/dx_k (e s_k) = e (s_k + pi c_k);  d2 = e (s_k + 2 pi c_k - pi^2 s_k)
        g.append(e * (s[k] + PI * c[k]) * o[0] * o[1])
        lap = lap + e * (s[k] + 2 * PI * c[k] - PI * PI * s[k]) * o[0] * o[1]
    return np.stack(g, axis=-1), lap


def vel_c3(x):
    return np.stack([np.sin(2 * PI * x[..., 1]), np.sin(2 * PI * x[..., 2]), np.sin(2 * PI * x[..., 0])], axis=-1)


def f_c3(x, D):
    g, lap = grad_lap_c3(x)
    b = vel_c3(x)
    return np.einsum('...k,...k->...', b, g) - D * lap


def upwind_tau(m, b_nodes, D, ell):
    """Double-valued tau: |b.n| + D/ell on the upwind side, D/ell on the other
    (tests/regression/HDG/TestHDGConvectionDiffusionReactionSource.cpp:182-202)."""
    a = m.arrays()
    n = face_normals(m)
    bn = np.einsum('fjd,fd->fj', b_nodes[a['faces']], n)
    tau = np.empty(a['faces'].shape + (2,))
    base = D / ell
    tau[..., 0] = np.where(bn > 0, np.abs(bn) + base, base)
    tau[..., 1] = np.where(bn > 0, base, np.abs(bn) + base)
    return tau


class Workload:
    """Inputs of one HDG problem, shared by the GPU path and the oracle."""

    def __init__(self, name, dim, p, mesh, kind, tau, tau_sides=1, D=None, ndD=0, b=None, dirichlet=None,
                 f_fn=None, u_exact=None, h=None, r_fn=None):
        self.name, self.dim, self.p, self.mesh, self.kind = name, dim, p, mesh, kind
        self.tau, self.tau_sides, self.D, self.ndD, self.b = tau, tau_sides, D, ndD, b
        self.dirichlet, self.f_fn, self.u_exact, self.h = dirichlet, f_fn, u_exact, h
        self.r_fn = r_fn      # reaction r(x) (CDRS model only), or None
        self.r_ip = None      # the reaction table setup_gpu uploaded

    def source_at(self, xip):
        return None if self.f_fn is None else np.ascontiguousarray(self.f_fn(xip))

    def reaction_at(self, xip):
        return None if self.r_fn is None else np.ascontiguousarray(self.r_fn(xip))


def diffusion_source_2d(n, p, jitter=0.1, seed=1234, name=None):
    """Unit square, n x n squares split in two, D = I, f = 2 pi^2 sin(pi x) sin(pi y),
    u = sin(pi x) sin(pi y), tau = 1 (the C1/C2 workloads)."""
    m = Mesh.structured(2, p, n, jitter, seed)
    tau = np.ones((m.n_faces, m.sizes['nNf']))
    return Workload(name or 'diffusion-source-2d-n%d-p%d' % (n, p), 2, p, m, 1, tau, 1,
                    dirichlet=dirichlet_from(m, u_sin2d), f_fn=f_sin2d, u_exact=u_sin2d, h=1.0 / n)


def convection_diffusion_3d(n, p, D=1e-2, jitter=0.05, seed=4321, name=None):
    """Unit cube, n^3 cubes split in six tetrahedra, D = 1e-2, b = (sin 2pi y,
    sin 2pi z, sin 2pi x), u = e^{x+y+z} sin sin sin, upwind double-valued tau
    with ell = h (the C3 workload)."""
    m = Mesh.structured(3, p, n, jitter, seed)
    x = m.arrays()['coords']
    b = vel_c3(x)
    h = 1.0 / n
    tau = upwind_tau(m, b, D, h)
    Dn = np.full(m.n_nodes, D)
    return Workload(name or 'convection-diffusion-3d-n%d-p%d' % (n, p), 3, p, m, 2, tau, 2, D=Dn, ndD=1, b=b,
                    dirichlet=dirichlet_from(m, u_c3), f_fn=lambda xx: f_c3(xx, D), u_exact=u_c3, h=h)


def setup_gpu(w, prob):
    """Feed a workload to an HDGProblem; returns the IP source table used."""
    prob.set_mesh(w.mesh)
    prob.set_model(w.kind)
    prob.set_fields(w.tau, w.tau_sides, w.D, w.ndD, w.b, w.dirichlet)
    xip = prob.ip_coords()
    f = w.source_at(xip)
    w.r_ip = w.reaction_at(xip)
    prob.set_source(f, w.r_ip)
    return f


def grf_2d(n, ell=0.1, sigma=1.0, seed=0):
    """Gaussian random field with squared-exponential covariance
    sigma^2 exp(-|x - y|^2 / (2 ell^2)) on the (n+1) x (n+1) vertices of the
    unit-square grid, sampled exactly by circulant embedding: the grid is
    embedded in a 2n x 2n torus of side 2, the covariance's eigenvalues are
    the FFT of its first row (negative round-off clipped), and
    Re FFT(sqrt(lambda / m^2) xi) with complex normal xi is a sample."""
    m = 2 * n
    d = np.minimum(np.arange(m), m - np.arange(m)) / n
    c = sigma * sigma * np.exp(-(d[:, None] ** 2 + d[None, :] ** 2) / (2.0 * ell * ell))
    lam = np.maximum(np.fft.fft2(c).real, 0.0)
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    g = np.fft.fft2(np.sqrt(lam / (m * m)) * xi).real
    return g[:n + 1, :n + 1]


def bilinear(grid, x):
    """Bilinear interpolation of vertex values grid[i, j] (at (i/n, j/n)) at points x."""
    n = grid.shape[0] - 1
    s = np.clip(x[..., 0] * n, 0.0, n - 1e-12)
    t = np.clip(x[..., 1] * n, 0.0, n - 1e-12)
    i, j = np.floor(s).astype(int), np.floor(t).astype(int)
    fs, ft = s - i, t - j
    return ((1 - fs) * (1 - ft) * grid[i, j] + fs * (1 - ft) * grid[i + 1, j]
            + (1 - fs) * ft * grid[i, j + 1] + fs * ft * grid[i + 1, j + 1])


def diffusion_grf_2d(n, p, seed, ell=0.1, sigma=1.0, jitter=0.1, mesh=None, mesh_seed=2024):
    """The learned-preconditioner workload: unit square, n x n squares split
    in two (jittered +-jitter h), diffusion coefficient k = exp(g) with g a
    GRF (ell, sigma) drawn from `seed`, source f = 1, u = 0 on the boundary,
    tau = 1.  k is a scalar node field (the reference's "DiffusionTensor"
    with one entry per node) interpolated bilinearly from the grid."""
    m = mesh if mesh is not None else Mesh.structured(2, p, n, jitter, mesh_seed)
    g = grf_2d(n, ell, sigma, seed)
    k = np.exp(bilinear(g, m.arrays()['coords']))
    tau = np.ones((m.n_faces, m.sizes['nNf']))
    return Workload('diffusion-grf-2d-n%d-p%d-s%d' % (n, p, seed), 2, p, m, 1, tau, 1, D=k, ndD=1,
                    dirichlet=np.zeros((m.n_faces, m.sizes['nNf'])), f_fn=lambda xx: np.ones(xx.shape[:-1]),
                    h=1.0 / n)


def fourier_source(modes):
    """f(x, y) = sum_m a_m sin(pi i_m x) sin(pi j_m y)."""
    def f(x):
        out = np.zeros(x.shape[:-1])
        for a, i, j in modes:
            out += a * np.sin(PI * i * x[..., 0]) * np.sin(PI * j * x[..., 1])
        return out
    return f


def batch_specs(count=64, seed=7, cells=(1000, 40000), orders=(1, 5), diff=(1e-2, 1e2)):
    """The batched workload: `count` independent 2-D diffusion problems with
    cell counts log-uniform in `cells` (n x n x 2 meshes, n = round(sqrt(c/2))),
    orders uniform in `orders`, scalar diffusivity log-uniform in `diff` and a
    source of 8 Fourier modes with N(0, 1) amplitudes, all from `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        c = math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1])))
        n = max(2, int(round(math.sqrt(c / 2.0))))
        p = int(rng.integers(orders[0], orders[1] + 1))
        D = math.exp(rng.uniform(math.log(diff[0]), math.log(diff[1])))
        modes = [(float(rng.standard_normal()), int(rng.integers(1, 5)), int(rng.integers(1, 5))) for _ in range(8)]
        out.append(dict(n=n, p=p, D=D, modes=modes, mesh_seed=seed * 1000 + k))
    return out


def batch_workload(spec, jitter=0.1):
    m = Mesh.structured(2, spec['p'], spec['n'], jitter, spec['mesh_seed'])
    nNf = m.sizes['nNf']
    return Workload('batch-n%d-p%d' % (spec['n'], spec['p']), 2, spec['p'], m, 1, np.ones((m.n_faces, nNf)), 1,
                    D=np.full(m.n_nodes, spec['D']), ndD=1, dirichlet=np.zeros((m.n_faces, nNf)),
                    f_fn=fourier_source(spec['modes']), h=1.0 / spec['n'])
