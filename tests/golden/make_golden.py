"""Writes the golden FFT fixtures in tests/golden/*.npz.

Each fixture holds an input and numpy's pocketfft FP64 result for one
transform of the one-rank decomp_2d_fft_3d path, with the reference's
conventions: Fortran (i fastest) layout, forward sign -1, no normalisation
(src/decomp_2d_constants.f90:86-87; examples/fft_physical_x/fft_c2c_x.f90:113),
r2c halving x for PHYSICAL_IN_X and z for PHYSICAL_IN_Z
(src/fft_common.f90:210-216), and c2r ignoring the imaginary parts of the
zero and Nyquist modes of the halved axis (src/fft_generic.f90:320-337).

Inputs are the reference tests' own analytic fields
(examples/fft_physical_x/fft_c2c_x.f90:64-73, fft_r2c_x.f90:66-73,
examples/fft_multiple_grids/fft_multiple_grids_reference.f90:15-46) and
uniform[-1,1) draws from numpy's PCG64 with a fixed seed, stored in the file.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# (nx, ny, nz): the reference base grid, even/odd neighbours, the multi-grid
# test grids, a power-of-two grid and degenerate axes.
SHAPES = [
    (17, 13, 11),
    (16, 12, 10),
    (18, 14, 12),
    (17, 15, 27),
    (33, 15, 11),
    (32, 16, 8),
    (9, 8, 7),
    (5, 1, 3),
    (1, 4, 2),
]


def analytic_real(nx, ny, nz):
    i = np.arange(1, nx + 1)[:, None, None] / nx
    j = np.arange(1, ny + 1)[None, :, None] / ny
    k = np.arange(1, nz + 1)[None, None, :] / nz
    return np.asfortranarray(i * j * k)


def fwd_c2c(x):
    return np.asfortranarray(np.fft.fftn(x))


def bwd_c2c(x):
    return np.asfortranarray(np.fft.ifftn(x, norm="forward"))


def r2c(x, fmt):
    axes = (2, 1, 0) if fmt == 1 else (0, 1, 2)
    return np.asfortranarray(np.fft.rfftn(x, axes=axes))


def c2r(c, shape, fmt):
    nx, ny, nz = shape
    if fmt == 1:
        return np.asfortranarray(np.fft.irfftn(c, s=(nz, ny, nx), axes=(2, 1, 0), norm="forward"))
    return np.asfortranarray(np.fft.irfftn(c, s=(nx, ny, nz), axes=(0, 1, 2), norm="forward"))


def spec_shape(shape, fmt):
    nx, ny, nz = shape
    return (nx // 2 + 1, ny, nz) if fmt == 1 else (nx, ny, nz // 2 + 1)


def main():
    rng = np.random.default_rng(20240601)
    for shape in SHAPES:
        nx, ny, nz = shape
        tag = "%dx%dx%d" % shape
        r = analytic_real(nx, ny, nz)
        # c2c: reference single-grid input re = im = r, and multi-grid
        # input re = r, im = -3r; plus a uniform random field.
        c_ref = np.asfortranarray(r + 1j * r)
        c_mg = np.asfortranarray(r - 3j * r)
        c_rand = np.asfortranarray(rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
        x_rand = np.asfortranarray(rng.uniform(-1, 1, shape))
        data = {
            "shape": np.array(shape, np.int32),
            "c_ref": c_ref, "c_ref_fwd": fwd_c2c(c_ref),
            "c_mg": c_mg, "c_mg_fwd": fwd_c2c(c_mg),
            "c_rand": c_rand, "c_rand_fwd": fwd_c2c(c_rand), "c_rand_bwd": bwd_c2c(c_rand),
            "r_ref": r, "r_rand": x_rand,
        }
        for fmt, fname in ((1, "x"), (3, "z")):
            data["r_ref_r2c_" + fname] = r2c(r, fmt)
            data["r_rand_r2c_" + fname] = r2c(x_rand, fmt)
            ss = spec_shape(shape, fmt)
            # a non-Hermitian spectrum: pins the c2r treatment of the
            # zero/Nyquist imaginary parts and of the 3D symmetry.
            s_rand = np.asfortranarray(rng.uniform(-1, 1, ss) + 1j * rng.uniform(-1, 1, ss))
            data["s_rand_" + fname] = s_rand
            data["s_rand_c2r_" + fname] = c2r(s_rand, shape, fmt)
        np.savez_compressed(os.path.join(HERE, "fft_%s.npz" % tag), **data)
        print("wrote fft_%s.npz" % tag)


if __name__ == "__main__":
    main()
