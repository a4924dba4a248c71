"""genspec (src/genspec.f90): the energy spectrum of one record of a field file.

  python -m ps3d_amd.genspec --filename <base>_fields.nc [--step k]

Reads the record's velocity (read_netcdf_fields, src/fields/field_netcdf.f90:276-433;
--step counts from 1, the default -1 takes the last record), forms ps3d's spectrum
with the library's transforms (fftxyp2s, dct of u and v, dst of w; genspec.f90:50-60),
bins the squared amplitudes in shells of K = sqrt(kx^2 + ky^2 + kz^2), normalises the
spectrum to the domain-averaged kinetic energy (genspec.f90:62-124) and writes
<filename without extension>_spectrum.asc with genspec's header (genspec.f90:140-165).
An existing spectrum file is not overwritten.
"""
import argparse
import os
import sys

import numpy as np

from . import closure, ncio
from . import gpu as G


def kinetic_energy(vel):
    """get_kinetic_energy (src/fields/field_diagnostics.f90:85-122): half the mean of
    |u|^2 with the lid levels weighted 1/2, over nx ny nz cells."""
    e = np.sum(vel ** 2, axis=0)                     # (nx, ny, nz+1)
    nx, ny, nz1 = e.shape
    tot = 0.5 * e[:, :, 1:-1].sum() + 0.25 * (e[:, :, 0].sum() + e[:, :, -1].sum())
    return tot / (nx * ny * (nz1 - 1))


def spectrum_of_file(filename, step=-1):
    """(shell wavenumbers, spectrum) of record `step` of a field file."""
    ncells, extent, origin = ncio.grid(filename)
    nx, ny, nz = ncells
    _, _, vel = ncio.read_fields(filename, nx, ny, nz, step=step, velocity=True)
    p = G.make_params(nx, ny, nz, lower=origin, extent=extent)
    with G.Ps3dGpu(p) as solver:
        return closure.velocity_spectrum(solver, vel, kinetic_energy(vel))


def write_spectrum(path, k, spec):
    with open(path, "w") as f:
        f.write(" # The power spectrum.\n")
        f.write(" # The first column is the wavenumber, the second column the spectrum.\n")
        f.write(" #         k   P(k)\n")
        for kk, s in zip(k, spec):
            f.write("  %24.16E  %24.16E\n" % (kk, s))


def main(argv=None):
    ap = argparse.ArgumentParser(description="Energy spectrum of a ps3d field file record.")
    ap.add_argument("--filename", required=True)
    ap.add_argument("--step", type=int, default=-1)
    a = ap.parse_args(argv)
    stem, ext = os.path.splitext(a.filename)
    if not ext:
        sys.exit("Error in reading the filename. File extension not found.")
    out = stem + "_spectrum.asc"
    if os.path.exists(out):
        sys.exit("Error: File '%s' already exists." % out)
    k, spec = spectrum_of_file(a.filename, a.step)
    write_spectrum(out, k, spec)
    return out


if __name__ == "__main__":
    main()
