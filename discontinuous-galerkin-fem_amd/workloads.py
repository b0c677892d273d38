"""Synthetic initial states of the benchmark workloads, from their closed forms.

* isentropic vortex: rho_inf = 1, u_inf = 1, v_inf = 0, p_inf = 1, strength beta at (xc, yc):
    r^2 = (x-xc)^2 + (y-yc)^2,  du = -beta/(2 pi) e^{(1-r^2)/2} (y-yc),
    dv = beta/(2 pi) e^{(1-r^2)/2} (x-xc),  T = 1 - (gamma-1) beta^2/(8 gamma pi^2) e^{1-r^2},
    rho = T^{1/(gamma-1)},  p = rho^gamma;  exact solution = IC(x - t, y) wrapped periodically.
  (the reference's example drives the same vortex in another scaling,
   examples/euler_vortex_test.cpp:24-49)
* per-element density noise rho <- rho (1 + 0.01 xi_e), xi_e ~ U(-1,1) from
  splitmix64(seed ^ e) (the same values on every device and host)
* Taylor-Green vortex: u = U0 sin x cos y, v = -U0 cos x sin y, rho = rho0,
  p = p0 + rho0 U0^2/4 (cos 2x + cos 2y), p0 = rho0 c0^2/gamma, c0 = U0/Mach.
States are conserved (rho, rho u, rho v, E) at the given points, shape [..., 4].
"""
import numpy as np


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed, k):
    z = splitmix64(np.uint64(seed) ^ np.asarray(k, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def density_noise(seed, K, Nq, amplitude=0.01):
    return np.repeat(amplitude * uniform_pm1(seed, np.arange(K))[:, None], Nq, axis=1)


def vortex(xy, beta=5.0, xc=5.0, yc=5.0, gamma=1.4, t=0.0, length=10.0, x0=0.0, rho_pert=None):
    x = x0 + np.mod(xy[..., 0] - t - x0, length)
    y = xy[..., 1]
    dx, dy = x - xc, y - yc
    r2 = dx * dx + dy * dy
    f = np.exp(0.5 * (1.0 - r2))
    u = 1.0 - beta / (2 * np.pi) * f * dy
    v = beta / (2 * np.pi) * f * dx
    T = 1.0 - (gamma - 1.0) * beta * beta / (8 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    if rho_pert is not None:
        rho = rho * (1.0 + rho_pert)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def _unit(z):
    return (np.asarray(z, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0


def batch_instances(n_instances=64, seed=5, sizes=(16, 32, 64, 128), orders=(1, 2, 3, 4)):
    """Instance parameters of the batched vortex workload, each drawn from splitmix64(seed ^ (8 i + k)):
    mesh size N uniform in `sizes`, order p uniform in `orders`, beta ~ U[3,6], centre ~ U[3,7]^2."""
    out = []
    for i in range(n_instances):
        r = splitmix64(np.uint64(seed) ^ np.arange(8 * i, 8 * i + 5, dtype=np.uint64))
        out.append({"n": int(sizes[int(r[0] % np.uint64(len(sizes)))]),
                    "p": int(orders[int(r[1] % np.uint64(len(orders)))]),
                    "beta": 3.0 + 3.0 * float(_unit(r[2])),
                    "xc": 3.0 + 4.0 * float(_unit(r[3])),
                    "yc": 3.0 + 4.0 * float(_unit(r[4]))})
    return out


def taylor_green(xy, gamma=1.4, mach=0.1, rho0=1.0, u0=1.0):
    x, y = xy[..., 0], xy[..., 1]
    c0 = u0 / mach
    p0 = rho0 * c0 * c0 / gamma
    u = u0 * np.sin(x) * np.cos(y)
    v = -u0 * np.cos(x) * np.sin(y)
    p = p0 + rho0 * u0 * u0 / 4.0 * (np.cos(2 * x) + np.cos(2 * y))
    rho = np.full_like(x, rho0)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)
