"""genspec (src/genspec.f90): the energy spectrum of one record of a field file.

  python -m ps3d_amd.genspec --filename <base>_fields.nc [--step k]

Reads the record's velocity (read_netcdf_fields, src/fields/field_netcdf.f90:276-433;
--step counts from 1, the default -1 takes the last record), forms ps3d's spectrum
with the library's transforms (fftxyp2s, dct of u and v, dst of w; genspec.f90:50-60),
bins the squared amplitudes in shells of K = sqrt(kx^2 + ky^2 + kz^2) and normalises the
spectrum to the domain-averaged kinetic energy (genspec.f90:62-124), all on the GPU
(ps3d_gpu_spectrum), and writes
<filename without extension>_spectrum.asc with genspec's header (genspec.f90:140-165).
An existing spectrum file is not overwritten.
"""
import argparse
import os
import sys


from . import ncio
from . import gpu as G


def spectrum_of_file(filename, step=-1):
    """(shell wavenumbers, spectrum) of record `step` of a field file."""
    ncells, extent, origin = ncio.grid(filename)
    nx, ny, nz = ncells
    _, _, vel = ncio.read_fields(filename, nx, ny, nz, step=step, velocity=True)
    p = G.make_params(nx, ny, nz, lower=origin, extent=extent)
    with G.Ps3dGpu(p) as solver:
        k, spec, _ = solver.spectrum(vel)
        return k, spec


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
