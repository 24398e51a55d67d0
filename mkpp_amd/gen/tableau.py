"""Rosenbrock tableaux (autonomous form) used by the integrator.

These are the published coefficient sets of Sandu et al., "Benchmarking stiff
ODE solvers for atmospheric chemistry problems II: Rosenbrock solvers",
Atmos. Environ. 31 (1997), as shipped by KPP.  The reference carries the same
numbers in ``src/mkpp/rosenbrock.py:27-133``; a published table has to be
identical, so the values are the same while the layout here is the build's
own.  A and C are stored row-wise strictly lower triangular:
``A[i][j]`` (0-based stage i > j) is ``a[i*(i-1)//2 + j]``.

Only ``gamma[0]`` enters W = I/(gamma h) - J (the reference integrates the
autonomous form, ``integrate.j2:76-86``).  ``new_f[s]`` marks stages that
evaluate F at their stage state; otherwise the previous F is reused.
"""

from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class Tableau:
    name: str
    stages: int
    a: tuple
    c: tuple
    m: tuple
    e: tuple
    alpha: tuple
    gamma: tuple
    new_f: tuple
    elo: float

    def A(self, i: int, j: int) -> float:
        return self.a[i * (i - 1) // 2 + j]

    def C(self, i: int, j: int) -> float:
        return self.c[i * (i - 1) // 2 + j]

    @property
    def n_f(self) -> int:
        return sum(1 for f in self.new_f if f)


_g2 = 1.0 + 1.0 / math.sqrt(2.0)

TABLEAUX = {
    "ros2": Tableau(
        "ROS-2", 2,
        a=(1.7071067811865475244,),
        c=(-1.1715728752538099024,),
        m=(0.87867965644035742774, 0.29289321881345247560),
        e=(0.29289321881345247560, 0.29289321881345247560),
        alpha=(0.0, 1.0),
        gamma=(1.7071067811865475244, -1.7071067811865475244),
        new_f=(True, True),
        elo=2.0,
    ),
    "ros3": Tableau(
        "ROS-3", 3,
        a=(1.0, 1.0, 0.0),
        c=(-1.0156171083877702092, 4.0759956452537699825, 9.2076794298330791242),
        m=(1.0, 6.1697947043828245593, -0.42772256543218573326),
        e=(0.5, -2.9079558716805469822, 0.22354069897811569627),
        alpha=(0.0, 0.43586652150845899942, 0.43586652150845899942),
        gamma=(0.43586652150845899942, 0.24291996454816804367, 2.1851380027664058512),
        new_f=(True, True, False),
        elo=3.0,
    ),
    "ros4": Tableau(
        "ROS-4", 4,
        a=(2.0, 1.867943637803922, 0.2344449711399156, 1.867943637803922, 0.2344449711399156, 0.0),
        c=(-7.137615036412310, 2.580708087951457, 0.6515950076447975,
           -2.137148994382534, -0.3214669691237626, -0.6949742501781779),
        m=(2.255570073418735, 0.2870493262186792, 0.4353179431840180, 1.093502252409163),
        e=(-0.2815431932141155, -0.07276199124938920, -0.1082196201495311, -1.093502252409163),
        alpha=(0.0, 1.145640000000000, 0.6552168638155900, 0.6552168638155900),
        gamma=(0.5728200000000000, -1.769193891319233, 0.7592633437920482, -0.1049021087100450),
        new_f=(True, True, True, False),
        elo=4.0,
    ),
    "rodas3": Tableau(
        "RODAS-3", 4,
        a=(0.0, 2.0, 0.0, 2.0, 0.0, 1.0),
        c=(4.0, 1.0, -1.0, 1.0, -1.0, -8.0 / 3.0),
        m=(2.0, 0.0, 1.0, 1.0),
        e=(0.0, 0.0, 0.0, 1.0),
        alpha=(0.0, 0.0, 1.0, 1.0),
        gamma=(0.5, 1.5, 0.0, 0.0),
        new_f=(True, False, True, True),
        elo=3.0,
    ),
    "rodas4": Tableau(
        "RODAS-4", 6,
        a=(1.544000000000000, 0.9466785280815826, 0.2557011698983284, 3.314825187068521,
           2.896124015972201, 0.9986419139977817, 1.221224509226641, 6.019134481288629,
           12.53708332932087, -0.6878860361058950, 1.221224509226641, 6.019134481288629,
           12.53708332932087, -0.6878860361058950, 1.0),
        c=(-5.668800000000000, -2.430093356833875, -0.2063599157091915, -0.1073529058151375,
           -9.594562251023355, -20.47028614809616, 7.496443313967647, -10.24680431464352,
           -33.99990352819905, 11.70890893206160, 8.083246795921522, -7.981132988064893,
           -31.52159432874371, 16.31930543123136, -6.058818238834054),
        m=(1.221224509226641, 6.019134481288629, 12.53708332932087, -0.6878860361058950, 1.0, 1.0),
        e=(0.0, 0.0, 0.0, 0.0, 0.0, 1.0),
        alpha=(0.0, 0.386, 0.210, 0.630, 1.0, 1.0),
        gamma=(0.25, -0.1043, 0.1035, -0.03620000000000023, 0.0, 0.0),
        new_f=(True, True, True, True, True, True),
        elo=4.0,
    ),
}

METHOD_IDS = {name: i for i, name in enumerate(["ros2", "ros3", "ros4", "rodas3", "rodas4"])}
