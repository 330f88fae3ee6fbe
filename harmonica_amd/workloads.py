"""
Seeded synthetic inputs of the benchmark and parity workloads.

Each generator returns plain numpy arrays (FP64).  Grids use inclusive
endpoints (numpy.linspace), as bordado.grid_coordinates does in the
reference's tests.

* terrain_layer_small  100 x 100 layer, 500 m prisms on a 50 km square,
  sinusoidal tops, g_z on a 64 x 64 grid at 1000 m (4.096e7 pairs).
* terrain_layer_large  1000 x 1000 layer, 100 m prisms on a 100 km square,
  fractal tops (spectral slope -2) in [0, 2000] m, bottoms at -5000 m, all
  ten fields on a 256 x 256 grid at 2500 m (6.55e10 pairs).
* batched_models       512 independent prism models with log-uniform prism
  (1e3-2e5) and observation (1e3-5e4) counts.
* terrain_magnetization  magnetisation of the terrain_layer_large prisms:
  induced by a 50,000 nT field (inclination -30, declination 12 degrees)
  with fractal susceptibility in [0.001, 0.05] SI, plus N(0, 0.2 A/m)
  remanence per component.
* global_tesseroid_layer  64,800 tesseroids of 1 x 1 degree over the globe
  (tops from a fractal surface in [-4, 4] km about the mean Earth radius,
  bottoms 30 km below it, densities 2670 +- 200 kg/m^3) and a 1-degree grid
  of 65,160 points 10 km above the mean radius.
* eqs_problem          60,000 scattered g_z data from 3 prisms + 1 mGal noise.
* inversion_layer      200 x 200 basement layer and a 128 x 128 grid at 500 m.
"""

import hashlib

import numpy as np


def sha256(workload):
    """SHA-256 (hex) of a generated input set: every array, scalar, tuple
    and list in ``workload`` (a dict, tuple or list, walked in key order),
    each array as its shape followed by its FP64 little-endian bytes."""
    h = hashlib.sha256()

    def feed(name, v):
        if isinstance(v, dict):
            for k in sorted(v):
                feed(f"{name}/{k}", v[k])
        elif isinstance(v, (tuple, list)):
            for i, item in enumerate(v):
                feed(f"{name}/{i}", item)
        else:
            a = np.ascontiguousarray(np.asarray(v, dtype="<f8"))
            h.update(f"{name}:{a.shape};".encode())
            h.update(a.tobytes())

    feed("", workload)
    return h.hexdigest()


def _grid(w, e, s, n, shape, height):
    east = np.linspace(w, e, shape[1])
    north = np.linspace(s, n, shape[0])
    E, N = np.meshgrid(east, north)
    return E, N, np.full_like(E, float(height))


def terrain_layer_small(seed=1):
    rng = np.random.default_rng(seed)
    centres = 250.0 + 500.0 * np.arange(100)
    E, N = np.meshgrid(centres, centres)
    lam = rng.uniform(5e3, 25e3, 4)
    theta = rng.uniform(0, 2 * np.pi, 4)
    phase = rng.uniform(0, 2 * np.pi, 4)
    waves = sum(
        np.sin(2 * np.pi * (E * np.cos(t) + N * np.sin(t)) / l + p)
        for l, t, p in zip(lam, theta, phase)
    )
    surface = 500.0 + 300.0 * waves
    density = 2670.0 + rng.normal(0, 50, surface.shape)
    return {
        "easting": centres.copy(),
        "northing": centres.copy(),
        "surface": surface,
        "reference": 0.0,
        "density": density,
        "coordinates": _grid(0, 50e3, 0, 50e3, (64, 64), 1000.0),
    }


def fractal_surface(n, rng, lo, hi, slope=-2.0):
    """White noise shaped to a power spectrum |k|^slope, min-max scaled."""
    noise = rng.normal(size=(n, n))
    spec = np.fft.rfft2(noise)
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx**2 + ky**2)
    k[0, 0] = 1.0
    spec *= k ** (slope / 2.0)
    spec[0, 0] = 0.0
    field = np.fft.irfft2(spec, s=(n, n))
    field = (field - field.min()) / (field.max() - field.min())
    return lo + (hi - lo) * field


def terrain_layer_large(seed=2, n_prisms_side=1000, n_obs_side=256):
    rng = np.random.default_rng(seed)
    spacing = 1e5 / n_prisms_side
    centres = spacing / 2 + spacing * np.arange(n_prisms_side)
    surface = fractal_surface(n_prisms_side, rng, 0.0, 2000.0)
    density = rng.uniform(2600.0, 2800.0, surface.shape)
    return {
        "easting": centres.copy(),
        "northing": centres.copy(),
        "surface": surface,
        "reference": -5000.0,
        "density": density,
        "coordinates": _grid(0, 1e5, 0, 1e5, (n_obs_side, n_obs_side), 2500.0),
    }


def terrain_magnetization(n_prisms_side=1000, seed=5):
    """(east, north, up) magnetisation in A/m of the terrain_layer_large
    prisms, in layer_to_prisms order (northing-major ravel)."""
    rng = np.random.default_rng(seed)
    chi = fractal_surface(n_prisms_side, rng, 0.001, 0.05).ravel()
    inc, dec = np.radians(-30.0), np.radians(12.0)
    h0 = 50000e-9 / (4e-7 * np.pi)  # |B0| / mu_0 in A/m
    unit = (np.cos(inc) * np.sin(dec), np.cos(inc) * np.cos(dec), -np.sin(inc))
    return tuple(chi * h0 * d + rng.normal(0.0, 0.2, chi.size) for d in unit)


def global_tesseroid_layer(seed=6, spacing=1.0):
    rng = np.random.default_rng(seed)
    mean_radius = 6371008.771415059  # WGS84 mean radius
    lon = np.arange(-180.0, 180.0, spacing)
    lat = np.arange(-90.0, 90.0, spacing)
    n_lon, n_lat = lon.size, lat.size
    tops = fractal_surface(max(n_lon, n_lat), rng, -4e3, 4e3)[:n_lat, :n_lon]
    W, S = np.meshgrid(lon, lat)
    tess = np.column_stack([W.ravel(), (W + spacing).ravel(), S.ravel(), (S + spacing).ravel(),
                            np.full(W.size, mean_radius - 3e4), (mean_radius + tops).ravel()])
    density = rng.uniform(2470.0, 2870.0, W.size)
    glon = np.linspace(-180.0, 179.0, 360)
    glat = np.linspace(-90.0, 90.0, 181)
    GL, GT = np.meshgrid(glon, glat)
    return {"tesseroids": tess, "density": density,
            "coordinates": (GL, GT, np.full(GL.shape, mean_radius + 1e4))}


def layer_to_prisms(easting, northing, surface, reference):
    """(m, 6) prisms of a layer in the reference's row order (northing-major
    meshgrid ravel, as DatasetAccessorPrismLayer._to_prisms)."""
    E, N = np.meshgrid(easting, northing)
    he = (easting[1] - easting[0]) / 2
    hn = (northing[1] - northing[0]) / 2
    ref = np.broadcast_to(np.asarray(reference, float), surface.shape)
    top = np.where(surface < ref, ref, surface)
    bottom = np.where(surface < ref, surface, ref)
    return np.column_stack([(E - he).ravel(), (E + he).ravel(), (N - hn).ravel(),
                            (N + hn).ravel(), bottom.ravel(), top.ravel()])


def batched_models(seed=3, n_models=512, m_range=(1e3, 2e5), n_range=(1e3, 5e4)):
    """List of (coordinates, prisms, density) for many uneven models."""
    rng = np.random.default_rng(seed)
    models = []
    for _ in range(n_models):
        m = int(round(np.exp(rng.uniform(np.log(m_range[0]), np.log(m_range[1])))))
        n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        sides = rng.uniform(50.0, 2000.0, (m, 3))
        centre = np.column_stack([rng.uniform(-10e3, 10e3, m), rng.uniform(-10e3, 10e3, m),
                                  rng.uniform(-20e3, 0.0, m)])
        top = np.minimum(centre[:, 2] + sides[:, 2] / 2, 0.0)
        prisms = np.column_stack([centre[:, 0] - sides[:, 0] / 2, centre[:, 0] + sides[:, 0] / 2,
                                  centre[:, 1] - sides[:, 1] / 2, centre[:, 1] + sides[:, 1] / 2,
                                  top - sides[:, 2], top])
        density = rng.normal(0.0, 300.0, m)
        coords = (rng.uniform(-10e3, 10e3, n), rng.uniform(-10e3, 10e3, n),
                  rng.uniform(100.0, 3000.0, n))
        models.append((coords, prisms, density))
    return models


def eqs_problem(seed=4, n_data=60000, grid_shape=(512, 512)):
    """Scattered data over a 100 km square, heights 500-1500 m, g_z of 3 random
    prisms (sides 2-10 km, tops 1-5 km deep, contrast +-200-400 kg/m^3);
    noise N(0, 1 mGal) is added by the caller after the forward model."""
    rng = np.random.default_rng(seed)
    coords = (rng.uniform(0, 1e5, n_data), rng.uniform(0, 1e5, n_data),
              rng.uniform(500.0, 1500.0, n_data))
    sides = rng.uniform(2e3, 10e3, (3, 3))
    centres = rng.uniform(0, 1e5, (3, 2))
    tops = -rng.uniform(1e3, 5e3, 3)
    prisms = np.column_stack([centres[:, 0] - sides[:, 0] / 2, centres[:, 0] + sides[:, 0] / 2,
                              centres[:, 1] - sides[:, 1] / 2, centres[:, 1] + sides[:, 1] / 2,
                              tops - sides[:, 2], tops])
    density = rng.choice([-1.0, 1.0], 3) * rng.uniform(200.0, 400.0, 3)
    noise = rng.normal(0.0, 1.0, n_data)
    return {
        "coordinates": coords,
        "prisms": prisms,
        "density": density,
        "noise": noise,
        "damping": 1e-3,
        "depth": 4500.0,
        "grid": _grid(0, 1e5, 0, 1e5, grid_shape, 1000.0),
    }


def inversion_layer(seed=5, n_side=200, n_obs_side=128):
    """Smooth basement (Gaussian-filtered noise, sigma 5 km) in [-6000, -1000] m,
    bottoms at -10000 m, contrast -300 + N(0, 20), observations at 500 m."""
    rng = np.random.default_rng(seed)
    spacing = 50e3 / n_side
    centres = spacing / 2 + spacing * np.arange(n_side)
    noise = rng.normal(size=(n_side, n_side))
    ky = np.fft.fftfreq(n_side, d=spacing)[:, None]
    kx = np.fft.rfftfreq(n_side, d=spacing)[None, :]
    sigma = 5e3
    gauss = np.exp(-2 * (np.pi * sigma) ** 2 * (kx**2 + ky**2))
    smooth = np.fft.irfft2(np.fft.rfft2(noise) * gauss, s=(n_side, n_side))
    smooth = (smooth - smooth.min()) / (smooth.max() - smooth.min())
    top = -6000.0 + 5000.0 * smooth
    density = -300.0 + rng.normal(0.0, 20.0, (n_side, n_side))
    return {
        "easting": centres.copy(),
        "northing": centres.copy(),
        "top": top,
        "bottom": np.full((n_side, n_side), -10000.0),
        "density": density,
        "coordinates": _grid(0, 50e3, 0, 50e3, (n_obs_side, n_obs_side), 500.0),
    }
