"""The reference's XML configuration file (CConfiguration.cpp:54-197).

``LbmConfiguration.from_xml`` reads the same elements with the same
semantics: numbers go through C ``atof`` / ``atoi`` (leading whitespace,
longest numeric prefix, 0 when there is none) and are then narrowed to the
arithmetic type T; the element ``domian-length`` keeps the reference's
spelling; ``subdomain-num`` is read as in the reference's MPI build
(``USE_MPI``, common.make:66-67). ``check()`` applies checkParameters
(CConfiguration.cpp:199-222), raising ``ValueError`` where the reference
asserts. ``subdomain(rank)`` derives a rank's size, origin and length as
CManager does (CManager.cpp:28-58), and ``solver_parameters`` runs the
CLbmSolver time-step adaptation for that subdomain.
"""
from __future__ import annotations

import math
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from .solver import compute_parameters

_WS = " \t\n\v\f\r"
_INT = re.compile(r"[+-]?\d+")
_DEC = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")
_HEX = re.compile(r"[+-]?0[xX]([0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)([pP][+-]?\d+)?")
_SPECIAL = re.compile(r"[+-]?(inf(inity)?|nan(\([0-9A-Za-z_]*\))?)", re.IGNORECASE)


def atoi(text) -> int:
    """C atoi: optional whitespace and sign, then digits; 0 without digits."""
    m = _INT.match((text or "").lstrip(_WS))
    return int(m.group(0)) if m else 0


def atof(text) -> float:
    """C atof (strtod): the longest decimal, hexadecimal, inf or nan prefix."""
    s = (text or "").lstrip(_WS)
    m = _HEX.match(s)
    if m:
        body = m.group(0)
        sign = -1.0 if body.startswith("-") else 1.0
        body = body.lstrip("+-")
        if "p" not in body.lower():
            body += "p0"
        return sign * float.fromhex(body)
    m = _SPECIAL.match(s)
    if m:
        v = m.group(0).lower().lstrip("+-")
        sign = -1.0 if m.group(0).startswith("-") else 1.0
        return sign * (math.inf if v.startswith("inf") else math.nan)
    m = _DEC.match(s)
    return float(m.group(0)) if m else 0.0


def _t(precision):
    return np.float64 if precision == "f64" else np.float32


@dataclass
class LbmConfiguration:
    precision: str = "f32"
    velocity: tuple = (0.0, 0.0, 0.0)
    acceleration: tuple = (0.0, 0.0, 0.0)
    viscosity: float = -1.0
    max_velocity: float = 0.1
    max_acceleration: float = 0.1
    domain_size: tuple = (32, 32, 32)
    domain_length: tuple = (32.0, 32.0, 32.0)
    subdomains: tuple = (1, 1, 1)
    cpu_subdomain_ratio: float = 0.0
    loops: int = 100
    timestep: float = 1.0
    do_benchmark: bool = False
    do_logging: bool = False
    do_validation: bool = False
    do_visualization: bool = False
    benchmark_dir: str = "."
    logging_dir: str = "."
    validation_dir: str = "."
    visualization_dir: str = "."
    visualization_rate: int = 1
    cpu_blocks: tuple = ((32, 32, 32), (32, 32, 32))             # alpha, beta
    gpu_grids: tuple = ((32, 16, 1), (32, 8, 1), (32, 8, 1))      # init, alpha, beta
    extra: dict = field(default_factory=dict)

    @classmethod
    def from_xml(cls, path, precision="f32") -> "LbmConfiguration":
        T = _t(precision)
        try:
            root = ET.parse(path).getroot()
        except (OSError, ET.ParseError) as e:
            raise ValueError(f"Loading XML configuration file \"{path}\" failed: {e}") from e
        if root.tag != "lbm-configuration":
            raise ValueError(f"{path}: root element is <{root.tag}>, not <lbm-configuration>")

        def node(*names):
            el = root
            for n in names:
                nxt = el.find(n)
                if nxt is None:
                    raise ValueError(f"{path}: missing element {'/'.join(names)}")
                el = nxt
            return el

        def text(*names):
            return node(*names).text or ""

        def f(*names):                  # atof, then narrowed to T
            return float(T(atof(text(*names))))

        def vec_f(*names):
            return tuple(f(*names, c) for c in "xyz")

        def vec_i(*names):
            return tuple(atoi(text(*names, c)) for c in "xyz")

        c = cls(precision=precision)
        # interpretPhysiscsData (CConfiguration.cpp:54-71)
        c.velocity = vec_f("physics", "velocity")
        c.acceleration = vec_f("physics", "acceleration")
        c.viscosity = f("physics", "viscosity")
        c.max_velocity = f("physics", "max-velocity")
        c.max_acceleration = f("physics", "max-acceleration")
        # interpretGridData (:73-93)
        c.domain_size = vec_i("grid", "domain-size")
        c.domain_length = vec_f("grid", "domian-length")
        c.subdomains = vec_i("grid", "subdomain-num")
        c.cpu_subdomain_ratio = f("grid", "cpu-subdomain-ratio")
        # interpretSimulationData (:95-117)
        c.loops = atoi(text("simulation", "loops"))
        c.timestep = f("simulation", "timestep")
        c.do_benchmark = bool(atoi(text("simulation", "benchmark", "do")))
        c.do_logging = bool(atoi(text("simulation", "logging", "do")))
        c.do_validation = bool(atoi(text("simulation", "validation", "do")))
        c.do_visualization = bool(atoi(text("simulation", "visualization", "do")))
        c.benchmark_dir = text("simulation", "benchmark", "output-dir")
        c.logging_dir = text("simulation", "logging", "output-dir")
        c.validation_dir = text("simulation", "validation", "output-dir")
        c.visualization_dir = text("simulation", "visualization", "output-dir")
        c.visualization_rate = atoi(text("simulation", "visualization", "rate"))
        # interpretCPUData / interpretGPUData (:119-162)
        c.cpu_blocks = (vec_i("cpu", "block-configuration", "alpha-block-configuration"),
                        vec_i("cpu", "block-configuration", "beta-block-configuration"))
        c.gpu_grids = tuple(vec_i("gpu", "grid-configuration", f"{k}-grid-configuration")
                            for k in ("init", "alpha", "beta"))
        c.check()
        return c

    def check(self):
        """checkParameters (CConfiguration.cpp:199-222)."""
        T = _t(self.precision)
        L = [T(v) for v in self.domain_length]
        S = self.domain_size
        eps = np.finfo(T).eps
        conds = [
            (self.timestep > 0, "timestep > 0"),
            (self.max_velocity > 0, "max velocity > 0"),
            (self.max_acceleration > 0, "max acceleration > 0"),
            (all(s > 0 for s in S), "domain size > 0"),
            (all(v > 0 for v in L), "domain length > 0"),
            (all(n > 0 for n in self.subdomains), "subdomain count > 0"),
            (0 <= self.cpu_subdomain_ratio <= 1, "0 <= CPU subdomain ratio <= 1"),
            (all(n > 0 and S[i] % n == 0 for i, n in enumerate(self.subdomains)),
             "domain size divisible by the subdomain count"),
            (self.visualization_rate > 0, "visualization rate > 0"),
            (all(v > 0 for b in self.cpu_blocks for v in b), "CPU block sizes > 0"),
            (all(v > 0 for g in self.gpu_grids for v in g), "GPU grid sizes > 0"),
            (all(g[0] * g[1] * g[2] <= 1024 for g in self.gpu_grids), "GPU grid <= 1024 threads"),
        ]
        if all(s > 0 for s in S):
            dx = [L[i] / T(S[i]) for i in range(3)]
            conds.append((all(abs(dx[i] - dx[(i + 1) % 3]) < eps for i in range(3)),
                          "cubic cells (equal length / size on every axis)"))
        for ok, what in conds:
            if not ok:
                raise ValueError(f"configuration check failed: {what}")

    def fields(self) -> list:
        """Every parsed field in a fixed order (velocity, acceleration,
        viscosity, max velocity / acceleration, size, length, subdomains, CPU
        ratio, loops, timestep, the four switches, visualization rate, CPU
        blocks, GPU grids)."""
        out = list(self.velocity) + list(self.acceleration)
        out += [self.viscosity, self.max_velocity, self.max_acceleration]
        out += list(self.domain_size) + list(self.domain_length) + list(self.subdomains)
        out += [self.cpu_subdomain_ratio, self.loops, self.timestep]
        out += [float(self.do_benchmark), float(self.do_logging), float(self.do_validation),
                float(self.do_visualization), self.visualization_rate]
        for b in self.cpu_blocks:
            out += list(b)
        for g in self.gpu_grids:
            out += list(g)
        return [float(v) for v in out]

    @property
    def ranks(self) -> int:
        return self.subdomains[0] * self.subdomains[1] * self.subdomains[2]

    def subdomain(self, rank: int):
        """(size, origin in cells, length) of one rank's subdomain, x fastest
        (CManager.cpp:34-58)."""
        T = _t(self.precision)
        n = self.subdomains
        i = rank % n[0]
        j = (rank // n[0]) % n[1]
        k = rank // (n[0] * n[1])
        size = tuple(self.domain_size[a] // n[a] for a in range(3))
        origin = (i * size[0], j * size[1], k * size[2])
        length = tuple(float(T(self.domain_length[a]) / T(n[a])) for a in range(3))
        return size, origin, length

    def solver_parameters(self, rank: int = 0) -> dict:
        """The CLbmSolver constructor's adaptation for one rank (tau, dt,
        dimensionless lid velocity and gravity); raises LbmError where the
        reference exits."""
        size, _, length = self.subdomain(rank)
        return compute_parameters(length_x=length[0], size_x=size[0],
                                  length_max=max(self.domain_length), velocity=self.velocity,
                                  acceleration=self.acceleration, viscosity=self.viscosity,
                                  timestep=self.timestep, max_velocity=self.max_velocity,
                                  max_acceleration=self.max_acceleration,
                                  precision=self.precision)

    @property
    def store_outputs(self) -> bool:
        """storeDensities / storeVelocities (CLbmSolver.cpp:40-41)."""
        return self.do_validation or self.do_visualization
