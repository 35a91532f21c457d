"""Small seeded scenes shared by the GPU parity tests."""
import numpy as np

N2 = 2.449489742783178   # sqrt(6)


def q6(a):
    a = np.asarray(a, dtype=np.float64)
    return np.array([float("%.6f" % v) for v in a.ravel()]).reshape(a.shape)


def spiral(n, phase=0.0, h=12.0):
    t = np.arange(n) / max(n - 1, 1)
    r = 5 + 10 * t
    return q6(np.stack([r * np.cos(4 * np.pi * t + phase), r * np.sin(4 * np.pi * t + phase), np.full(n, h)], 1))


def bumps_dem(n=12, spacing=30.0, seed=3):
    rng = np.random.default_rng(seed)
    ax = q6(-(n - 1) * spacing / 2 + spacing * np.arange(n))
    X, Y = np.meshgrid(ax, ax)
    Z = 0.006 * X - 0.004 * Y
    for _ in range(3):
        cx, cy = rng.uniform(ax[0], ax[-1], 2)
        Z = Z + rng.uniform(0.5, 2) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * rng.uniform(40, 90) ** 2))
    return ax, ax.copy(), q6(Z)


class Tiny:
    """Handful of voxels and pulses; every check against the oracle is cheap."""

    def __init__(self, n=24, S=512, grid=(7, 5, 4), seed=9, fc=1e9, fs=800e6, dem=None, z0=0.0):
        self.rng = np.random.default_rng(seed)
        self.n, self.S = n, S
        self.tx, self.rx = spiral(n, h=z0 + 12.0), spiral(n, np.pi / 2, h=z0 + 12.0)
        self.dem = bumps_dem() if dem is None else dem
        self.gx = q6(19.85 + 0.07 * np.arange(grid[0]))
        self.gy = q6(-0.15 + 0.07 * np.arange(grid[1]))
        self.gz = q6(z0 - 5.2 + 0.13 * np.arange(grid[2]))
        self.fc, self.c, self.fs, self.n1, self.n2, self.up = fc, 3e8, fs, 1.0, N2, 4

    def params(self, bsar):
        return bsar.Params(self.fc, self.c, self.fs, self.n1, self.n2, self.up)

    def bp(self):
        return dict(c=self.c, fs4=self.fs * self.up, lam=self.c / self.fc, n1=self.n1, n2=self.n2)

    def random_root(self, ndata=None):
        nd = self.n if ndata is None else ndata
        return self.rng.normal(size=(nd, self.S)) + 1j * self.rng.normal(size=(nd, self.S))

    def target_root(self, oracle, pts=None, bw=300e6):
        if pts is None:
            pts = np.array([[self.gx[len(self.gx) // 2], self.gy[len(self.gy) // 2], self.gz[len(self.gz) // 2]]])
        sim = dict(fc=self.fc, c=self.c, fs=self.fs, f0=-bw / 2, f1=bw / 2, T=1e-7, fs_chirp=self.fs, n1=self.n1,
                   n2=self.n2, S=self.S)
        _, _, root = oracle.simulate(sim, self.dem, self.tx, self.rx, pts, np.ones(len(pts), complex))
        return np.ascontiguousarray(root.T)      # [pulse][range]


def peak_err(a, ref):
    """max |a - ref| / max |ref| (absolute max |a - ref| when ref is all zero)."""
    d = float(np.abs(a - ref).max()) if np.size(ref) else 0.0
    m = float(np.abs(ref).max()) if np.size(ref) else 0.0
    return d / m if m > 0 else d


def db(x):
    return 20 * np.log10(max(x, 1e-300))
