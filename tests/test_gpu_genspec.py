"""genspec (src/genspec.f90) on a field file: a velocity with a single Fourier /
cosine component puts all its energy into one shell, and the spectrum integrates
to the kinetic energy (genspec.f90:62-124)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_single_mode_spectrum(tmp_path):
    from ps3d_amd import genspec, ics, ncio
    from oracle import oracle as O
    n = 32
    ext = (2 * np.pi, 2 * np.pi, np.pi)
    x, y, z = ics.grid(n, n, n, (0.0, 0.0, 0.0), ext)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    vel = np.zeros((3, n, n, n + 1))
    vel[0] = np.cos(3 * X) * np.cos(4 * Z)          # |K| = 5 (kx = 3, kz = 4)
    ff = ncio.FieldFile(str(tmp_path / "mode"), n, n, n, (0.0, 0.0, 0.0), ext, derived=False)
    ff.write(0.0, vel, np.zeros_like(vel))
    print("genspec", flush=True)
    out = genspec.main(["--filename", ff.path])
    assert os.path.basename(out) == "mode_fields_spectrum.asc"
    data = np.loadtxt(out, comments="#")
    k, spec = data[:, 0], data[:, 1]
    dk = k[1] - k[0]
    ke = O.kinetic_energy(vel)
    assert abs(ke - 0.125) < 1e-14                  # mean of cos^2 cos^2 / 2
    assert abs(np.sum(spec) * dk - ke) < 1e-14
    peak = int(np.argmax(spec))
    assert peak == int(5 / dk)
    others = np.delete(spec, peak)
    assert np.max(np.abs(others)) < 1e-12 * spec[peak]
    with pytest.raises(SystemExit):                  # an existing spectrum is not overwritten
        genspec.main(["--filename", ff.path])
