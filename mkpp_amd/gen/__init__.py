"""Mechanism compiler: OpenAtmos v1 JSON -> mass-action model -> sparse LU plan
-> HIP/CDNA4 kernels, plan JSON and oracle tables.

Generation runs on the build host only.  Its products (``mkpp_amd/csrc/gen/*``
and ``mkpp_amd/plans/*``) are committed, so nothing here runs on a GPU host.
"""

from .model import Mechanism, load_mechanism  # noqa: F401
from .tableau import TABLEAUX  # noqa: F401
