"""The y transforms of PoissonFFT2D_* as their trigonometric sums (no FFT),
for lengths where a dense basis would not fit in memory: the basis is built
in blocks of modes, each angle reduced on integers to [0, 2 pi) before its
sine or cosine (an unreduced angle of order N^2 pi loses about 1e-12 at
N = 4096), so every entry is exact to a few units of FP64 rounding and the
blocked products (BLAS, FP64) add about sqrt(N) units.

Matrix row j = coefficient j (poisson_fft2d.cpp):
  DD (:36-44):   sqrt(2/(N+1)) sin(pi (j+1)(n+1) / (N+1))
  DN (:241-250): (-1)^j 2/sqrt(2N+1) sin(pi (n+1)(2(N-j)-1) / (2N+1))
  ND (:180-188): (-1)^j 2/sqrt(2N+1) cos(pi (2n+1)(2(N-j)-1) / (2(2N+1)))
  NN (:102-117): -sqrt(2/N) cos(pi (N-1-j)(2n+1) / 2N), j <= N-2; sqrt(1/N) for j = N-1
  PP (:367-403): sqrt(2/N) sin(2 pi q_j (n+1) / N + pi (2j+1) / 4), q_j = N/2 - ceil(j/2)
The inverse (transform_transpose) is the transpose."""
import numpy as np

D, N, P = 0, 1, 2
KINDS = {(D, D): "DD", (D, N): "DN", (N, D): "ND", (N, N): "NN", (P, P): "PP"}


def _block(kind, ny, j0, j1):
    j = np.arange(j0, j1, dtype=np.int64)[:, None]
    n = np.arange(ny, dtype=np.int64)[None, :]
    pi = np.pi
    sgn = np.where(j % 2 == 0, 1, -1)
    if kind == "DD":
        k = ((j + 1) * (n + 1)) % (2 * (ny + 1))
        return np.sqrt(2 / np.float64(ny + 1)) * np.sin(pi * k.astype(np.float64) / (ny + 1))
    if kind == "DN":
        k = ((n + 1) * (2 * (ny - j) - 1)) % (2 * (2 * ny + 1))
        return sgn * (2 / np.sqrt(np.float64(2 * ny + 1))) * np.sin(pi * k.astype(np.float64) / (2 * ny + 1))
    if kind == "ND":
        k = ((2 * n + 1) * (2 * (ny - j) - 1)) % (4 * (2 * ny + 1))
        return sgn * (2 / np.sqrt(np.float64(2 * ny + 1))) * np.cos(pi * k.astype(np.float64) /
                                                                      (2 * (2 * ny + 1)))
    if kind == "PP":
        # angle = pi (8 q (n+1) + N (2j+1)) / 4N
        q = ny // 2 - (j + 1) // 2
        k = (8 * q * (n + 1) + ny * (2 * j + 1)) % (8 * ny)
        return np.sqrt(2 / np.float64(ny)) * np.sin(pi * k.astype(np.float64) / (4 * ny))
    k = ((ny - 1 - j) * (2 * n + 1)) % (4 * ny)
    m = -np.sqrt(2 / np.float64(ny)) * np.cos(pi * k.astype(np.float64) / (2 * ny))
    m[(j[:, 0] == ny - 1)] = np.sqrt(1 / np.float64(ny))
    return m


def transform(bc_neg, bc_pos, x, inverse=False, block=1024):
    """Forward (inverse = False) or transposed transform of the rows of x."""
    kind = KINDS[(bc_neg, bc_pos)]
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    ny = x.shape[1]
    out = np.zeros(x.shape)
    for j0 in range(0, ny, block):
        j1 = min(ny, j0 + block)
        m = _block(kind, ny, j0, j1)
        if inverse:
            out += x[:, j0:j1] @ m
        else:
            out[:, j0:j1] = x @ m.T
    return out
