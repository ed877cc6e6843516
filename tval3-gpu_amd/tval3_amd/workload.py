"""Synthetic TVAL3 workloads: phantoms, transducer geometry, noise.

Deterministic numpy generators (PCG64 via numpy.random.default_rng) shared
by the tests and the benchmark.  Volume layout is the reference's
column-major order: U[y + x*p + z*p*q] with p = dim_y (stride 1),
q = dim_x, r = dim_z (3d/container_device.h:21-23; voxel index of
3d/tval3_gpu_3d_kernels.cu:1501).

Conventions
-----------
* The volume occupies the world box [-1, 1]^3 (2-D: [-1, 1]^2); voxel size
  h = 2/N per axis; voxel (x, y, z) spans origin + [x, x+1)*h etc.
* Transducers lie uniformly on a sphere of radius 1.1 centred on the box.
* Ray order is emitter-major over all E*R pairs (the reference's compacted
  order with every path enabled, 3d/mat_vec_mul_dynamic.h:501-545).
* Seeds: phantom 1, geometry 2, noise 3, dense A 4.
"""
import numpy as np

SEED_PHANTOM, SEED_GEOMETRY, SEED_NOISE, SEED_DENSE = 1, 2, 3, 4
BACKGROUND = 1.0 / 1500.0


def _centres(N):
    return -1.0 + (np.arange(N) + 0.5) * (2.0 / N)


def phantom_2d(N=64, n_ellipses=8, seed=SEED_PHANTOM):
    """Piecewise-constant 2-D phantom of n_ellipses ellipses on [-1,1]^2.

    Ellipses are painted in order (the last one wins); centre U[-0.7,0.7]^2,
    semi-axes U[0.05,0.4], angle U[0,pi), value U[0,1].  Returns the
    column-major flattened image (p = q = N) as float64.
    """
    rng = np.random.default_rng(seed)
    c = _centres(N)
    X, Y = np.meshgrid(c, c, indexing="xy")     # Y[i, j] = c[i] (row = y), X[i, j] = c[j]
    img = np.zeros((N, N))                        # img[y, x]
    for _ in range(n_ellipses):
        cx, cy = rng.uniform(-0.7, 0.7, 2)
        a, b = rng.uniform(0.05, 0.4, 2)
        th = rng.uniform(0.0, np.pi)
        v = rng.uniform(0.0, 1.0)
        xr = (X - cx) * np.cos(th) + (Y - cy) * np.sin(th)
        yr = -(X - cx) * np.sin(th) + (Y - cy) * np.cos(th)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] = v
    # column-major with y fastest: index y + x*p
    return np.ascontiguousarray(img.T).reshape(-1).copy()


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def phantom_3d(N=64, n_ellipsoids=6, seed=SEED_PHANTOM, contrast=0.03, dims=None):
    """Slowness phantom: background 1/1500 s/m plus n_ellipsoids rotated
    ellipsoids at (1/1500)(1 +/- contrast), centre U[-0.6,0.6]^3, semi-axes
    U[0.08,0.35].  Returns float64 array of length p*q*r, column-major.
    dims = (p, q, r) overrides the cubic N."""
    p, q, r = dims if dims is not None else (N, N, N)
    rng = np.random.default_rng(seed)
    cy, cx, cz = _centres(p), _centres(q), _centres(r)
    # array indexed [z, x, y] so that ravel() gives y fastest, then x, then z
    Z, X, Y = np.meshgrid(cz, cx, cy, indexing="ij")
    vol = np.full((r, q, p), BACKGROUND)
    for _ in range(n_ellipsoids):
        c = rng.uniform(-0.6, 0.6, 3)
        ax = rng.uniform(0.08, 0.35, 3)
        Rm = _random_rotation(rng)
        sign = 1.0 if rng.uniform() < 0.5 else -1.0
        val = BACKGROUND * (1.0 + sign * contrast)
        d = np.stack([X - c[0], Y - c[1], Z - c[2]], axis=-1) @ Rm    # body frame
        inside = ((d / ax) ** 2).sum(-1) <= 1.0
        vol[inside] = val
    return vol.reshape(-1).copy()


def sphere_points(k, radius=1.1, seed=SEED_GEOMETRY):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(k, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return radius * v


def transducer_geometry(E, R, N=None, dims=None, radius=1.1, seed=SEED_GEOMETRY):
    """Emitters and receivers uniform on a sphere around the [-1,1]^3 box.

    Returns a dict usable by both the Siddon and Bresenham operators:
    e_pos/r_pos world coordinates (x, y, z), e_vox/r_vox the voxel
    coordinates (x, y, z) of the nearest voxel inside the grid."""
    p, q, r = dims if dims is not None else (N, N, N)
    pts = sphere_points(E + R, radius, seed)
    e_pos, r_pos = pts[:E], pts[E:]
    nx, ny, nz = q, p, r
    voxel = (2.0 / nx, 2.0 / ny, 2.0 / nz)
    origin = (-1.0, -1.0, -1.0)

    def vox(P):
        out = np.empty_like(P, dtype=np.int32)
        for a, n in enumerate((nx, ny, nz)):
            out[:, a] = np.clip(np.floor((P[:, a] - origin[a]) / voxel[a]), 0, n - 1)
        return out

    return dict(E=E, R=R, mps=1, rv=(nx, ny, nz), origin=origin, voxel=voxel,
                e_pos=e_pos, r_pos=r_pos, e_vox=vox(e_pos), r_vox=vox(r_pos),
                scale=voxel[0], use_path=None)


def dense_gaussian(m, n, seed=SEED_DENSE):
    """i.i.d. N(0, 1/m) entries, column-major (Fortran) float64."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.normal(0.0, 1.0 / np.sqrt(m), size=(m, n)))


def add_noise(Au, level, seed=SEED_NOISE):
    """b = Au + level*||Au|| xi/||xi||, xi ~ N(0, I)."""
    rng = np.random.default_rng(seed)
    xi = rng.normal(size=Au.shape)
    return Au + level * np.linalg.norm(Au) * xi / np.linalg.norm(xi)
