"""Synthetic ProCell inputs with the shapes of the standard workloads.

Each workload is a fluorescence histogram (value, count rows in file order),
a cell-type table (proportion, mean, sd in file order; mean < 0 is
quiescent), t_max and phi (0 means "smallest value with a nonzero count",
the README rule for an absent -p).

Values are computed with Python floats (libm pow/exp), counts with numpy's
PCG64 Poisson sampler, so a given seed yields the same bytes on every host
of this image.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    values: np.ndarray          # float64 [rows]
    counts: np.ndarray          # uint64 [rows]
    proportion: np.ndarray      # float64 [types]
    mean: np.ndarray            # float64 [types]
    sd: np.ndarray              # float64 [types]
    t_max: float
    phi: float                  # 0: auto
    track_ratio: bool = True
    meta: dict = field(default_factory=dict)

    @property
    def n_cells(self) -> int:
        return int(self.counts.sum())


def _lognormal_counts(values, total, mu, sigma, rng):
    lam = [math.exp(-((math.log(v) - mu) ** 2) / (2.0 * sigma * sigma)) for v in values]
    s = sum(lam)
    lam = np.array([total * x / s for x in lam], dtype=np.float64)
    return rng.poisson(lam).astype(np.uint64)


def small_histogram(seed: int = 0) -> Workload:
    """64 log-spaced bins from 1 to 1e4, ~1e5 cells, the README's three
    types (0.53 48.33 21.6 / 0.29 86.3 26.8 / 0.18 -1 -1), t_max 168,
    phi = smallest populated bin."""
    rng = np.random.default_rng([seed, 1])
    values = [10.0 ** (4.0 * i / 63.0) for i in range(64)]
    counts = _lognormal_counts(values, 1e5, math.log(800.0), 1.0, rng)
    return Workload(
        name="hist64_1e5_types3",
        values=np.array(values, dtype=np.float64),
        counts=counts,
        proportion=np.array([0.53, 0.29, 0.18]),
        mean=np.array([48.33, 86.3, -1.0]),
        sd=np.array([21.6, 26.8, -1.0]),
        t_max=168.0,
        phi=0.0,
        track_ratio=True,
        meta={"bins": 64, "target_cells": 1e5},
    )


def large_histogram(seed: int = 0) -> Workload:
    """1024 log-spaced bins from 1 to 1e5, ~1e7 cells, 4 proliferating types
    (means U(30,100), sd = U(0.2,0.4) x mean) plus one quiescent type, ratios
    Dirichlet(1,...,1), t_max 240, phi 50, per-type tracking."""
    rng = np.random.default_rng([seed, 2])
    values = [10.0 ** (5.0 * i / 1023.0) for i in range(1024)]
    counts = _lognormal_counts(values, 1e7, math.log(800.0), 1.0, rng)
    means = rng.uniform(30.0, 100.0, size=4)
    sds = rng.uniform(0.2, 0.4, size=4) * means
    ratios = rng.dirichlet(np.ones(5))
    return Workload(
        name="hist1024_1e7_types5",
        values=np.array(values, dtype=np.float64),
        counts=counts,
        proportion=ratios.astype(np.float64),
        mean=np.concatenate([means, [-1.0]]).astype(np.float64),
        sd=np.concatenate([sds, [-1.0]]).astype(np.float64),
        t_max=240.0,
        phi=50.0,
        track_ratio=True,
        meta={"bins": 1024, "target_cells": 1e7},
    )


def parameter_sets(seed: int = 0, n_sets: int = 512, scale: float = 1.0) -> list:
    """Heterogeneous parameter sets of the batched plain method.  Per set:
    128-1024 log-spaced bins from 1 to 1e5 with lognormal-shaped Poisson
    counts totalling 10^x cells, x ~ N(4.5, 0.75) clipped to [3, 6] (times
    `scale`); 1-5 types with means U(24,120), sd = U(0.15,0.5) x mean and,
    for two or more types, a quiescent last type half of the time; ratios
    Dirichlet(1,...,1); t_max U(72,336); phi U(1,100); per-type tracking."""
    rng = np.random.default_rng([seed, 3])
    sets = []
    for i in range(n_sets):
        bins = int(rng.integers(128, 1025))
        cells = 10.0 ** float(np.clip(rng.normal(4.5, 0.75), 3.0, 6.0)) * scale
        values = [10.0 ** (5.0 * j / (bins - 1)) for j in range(bins)]
        counts = _lognormal_counts(values, cells, math.log(800.0), 1.0, rng)
        n_types = int(rng.integers(1, 6))
        quiescent = n_types >= 2 and bool(rng.integers(0, 2))
        n_prolif = n_types - (1 if quiescent else 0)
        means = rng.uniform(24.0, 120.0, size=n_prolif)
        sds = rng.uniform(0.15, 0.5, size=n_prolif) * means
        if quiescent:
            means = np.concatenate([means, [-1.0]])
            sds = np.concatenate([sds, [-1.0]])
        ratios = rng.dirichlet(np.ones(n_types))
        sets.append(Workload(
            name=f"set{i}_bins{bins}",
            values=np.array(values, dtype=np.float64),
            counts=counts,
            proportion=ratios.astype(np.float64),
            mean=means.astype(np.float64),
            sd=sds.astype(np.float64),
            t_max=float(rng.uniform(72.0, 336.0)),
            phi=float(rng.uniform(1.0, 100.0)),
            track_ratio=True,
            meta={"bins": bins, "target_cells": cells},
        ))
    return sets


def calibration_instance(seed: int = 0) -> Workload:
    """Single-instance calibration problem: 256 log-spaced bins from 1 to 1e4
    with ~1e6 cells (lognormal-shaped Poisson counts, mu = ln 800, sigma 1),
    hidden types drawn from the seed -- two proliferating types with means
    U(30,100) and sd = U(0.2,0.4) x mean plus one quiescent type, ratios
    Dirichlet(1,1,1) -- t_max 168, phi = smallest populated bin.  The
    relaxed estimator's temperature is beta = 0.5 (meta["beta"])."""
    rng = np.random.default_rng([seed, 4])
    values = [10.0 ** (4.0 * i / 255.0) for i in range(256)]
    counts = _lognormal_counts(values, 1e6, math.log(800.0), 1.0, rng)
    means = rng.uniform(30.0, 100.0, size=2)
    sds = rng.uniform(0.2, 0.4, size=2) * means
    ratios = rng.dirichlet(np.ones(3))
    return Workload(
        name="hist256_1e6_types3_calibration",
        values=np.array(values, dtype=np.float64),
        counts=counts,
        proportion=ratios.astype(np.float64),
        mean=np.concatenate([means, [-1.0]]).astype(np.float64),
        sd=np.concatenate([sds, [-1.0]]).astype(np.float64),
        t_max=168.0,
        phi=0.0,
        track_ratio=True,
        meta={"bins": 256, "target_cells": 1e6, "beta": 0.5},
    )


def calibration_batch(seed: int = 0, n_sets: int = 256, cells: float = 2e5):
    """Batched calibration problem: n_sets candidate histograms, each of
    256 log-spaced bins from 1 to 1e4 with ~`cells` cells (lognormal-shaped
    Poisson counts, mu = ln 800, sigma 1), generating types drawn per set --
    1-3 proliferating types with means U(30,100) and sd = U(0.2,0.4) x mean,
    plus a quiescent type when that makes 2-4 types in all (always for one
    proliferating type, else half the time) -- ratios Dirichlet(2,...,2),
    t_max U(96,240), phi = smallest populated bin.  Returns (generating
    workloads, starting workloads): the start perturbs every mean and sd by
    a factor U(0.85,1.15) and mixes the ratios halfway towards uniform."""
    rng = np.random.default_rng([seed, 5])
    values = [10.0 ** (4.0 * i / 255.0) for i in range(256)]
    gen, start = [], []
    for i in range(n_sets):
        counts = _lognormal_counts(values, cells, math.log(800.0), 1.0, rng)
        n_prolif = int(rng.integers(1, 4))
        quiescent = n_prolif == 1 or bool(rng.integers(0, 2))
        means = rng.uniform(30.0, 100.0, size=n_prolif)
        sds = rng.uniform(0.2, 0.4, size=n_prolif) * means
        if quiescent:
            means = np.concatenate([means, [-1.0]])
            sds = np.concatenate([sds, [-1.0]])
        n_types = len(means)
        ratios = rng.dirichlet(np.full(n_types, 2.0))
        t_max = float(rng.uniform(96.0, 240.0))
        w = Workload(name=f"calib{i}", values=np.array(values, dtype=np.float64), counts=counts,
                     proportion=ratios.astype(np.float64), mean=means.astype(np.float64),
                     sd=sds.astype(np.float64), t_max=t_max, phi=0.0, track_ratio=True,
                     meta={"bins": 256, "target_cells": cells})
        f_m = rng.uniform(0.85, 1.15, size=n_types)
        f_s = rng.uniform(0.85, 1.15, size=n_types)
        prolif = means >= 0
        m0 = np.where(prolif, means * f_m, -1.0)
        s0 = np.where(prolif, sds * f_s, -1.0)
        p0 = 0.5 * ratios + 0.5 / n_types
        p0 = p0 / p0.sum()
        gen.append(w)
        start.append(Workload(name=f"calib{i}_start", values=w.values, counts=counts, proportion=p0,
                              mean=m0, sd=s0, t_max=t_max, phi=0.0, track_ratio=True, meta=dict(w.meta)))
    return gen, start


def scaled(w: Workload, factor: float, name: str | None = None) -> Workload:
    """Same shape with every count multiplied by `factor` (rounded down)."""
    counts = np.floor(w.counts.astype(np.float64) * factor).astype(np.uint64)
    return Workload(name or f"{w.name}_x{factor:g}", w.values.copy(), counts,
                    w.proportion.copy(), w.mean.copy(), w.sd.copy(), w.t_max,
                    w.phi, w.track_ratio, dict(w.meta, scale=factor))


def write_files(w: Workload, histogram_path: str, types_path: str) -> None:
    """Write the two reference input files (floats in repr form, which
    round-trips exactly)."""
    with open(histogram_path, "w") as f:
        for v, c in zip(w.values, w.counts):
            f.write(f"{float(v)!r} {int(c)}\n")
    with open(types_path, "w") as f:
        for p, m, s in zip(w.proportion, w.mean, w.sd):
            f.write(f"{float(p)!r} {float(m)!r} {float(s)!r}\n")
