"""An x-mirror-symmetric initial state (about the domain's centre line) for the
symmetry tests: density, pressure and v even in x, u odd, built on the left half
and mirrored so that the symmetry is exact in floating point."""
import numpy as np


def mirror_state(N_x, N_y, ng=2, gamma=1.4):
    """AoS state (N_y + 2 ng, N_x + 2 ng, 4) on [0, 1] x [0, N_y / N_x]."""
    M, N = N_x + 2 * ng, N_y + 2 * ng
    dx = 1.0 / N_x
    H = N_y / N_x
    x = (np.arange(M) - ng + 0.5) * dx
    y = (np.arange(N) - ng + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    rho = 1 + 0.5 * np.exp(-((X - 0.5) ** 2) / 0.01) * (1 + 0.3 * np.sin(2 * np.pi * Y / H))
    p = 1 + 0.8 * np.exp(-((X - 0.45) ** 2) / 0.004)
    u = 0.2 * np.sin(2 * np.pi * (X - 0.5)) + 0.1 * (X - 0.3)
    v = 0.1 * np.cos(2 * np.pi * Y / H)
    U = np.empty((N, M, 4))
    U[..., 0] = rho
    U[..., 1] = rho * u
    U[..., 2] = rho * v
    U[..., 3] = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    h = M // 2
    U[:, M - h:] = U[:, :h][:, ::-1]
    U[:, M - h:, 1] = -U[:, M - h:, 1]
    return U


def mirror_defects(U, M, N):
    """Number of values that differ in bits from the mirror image (u negated)."""
    U = np.asarray(U).reshape(N, M, 4)
    Um = U[:, ::-1].copy()
    Um[..., 1] = -Um[..., 1]
    return int((U.view(np.uint64) != Um.view(np.uint64)).sum())
