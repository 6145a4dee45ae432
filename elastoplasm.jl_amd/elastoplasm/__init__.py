"""ElastoPlasm.jl's explicit MPM step on MI355X: host-side mirror of the
reference API (``ic_slump``/``slump``/``slump_``, ``ic_collapse``/``collapse``,
``require``/``kwargser``, the ``Mesh``/``Point`` structs) over the C ABI of
libelastoplasm_hip.so."""
from .config import kwargser, require
from .ic import (NT, get_geom, get_slump, setup_cmpr, setup_mesh, setup_mpts, setup_time, get_collapse,
                 get_column, mesh_e2n)
from .types import Mesh, Point, Solid
from .workflow import (collapse, elastodynamic_, elastoplasm, elastoplastic_, ic_collapse, ic_ensemble, ic_slump,
                       slump, slump_)
from . import grf, hip

__all__ = ["kwargser", "require", "NT", "get_geom", "get_slump", "setup_cmpr", "setup_mesh", "setup_mpts",
           "setup_time", "get_collapse", "get_column", "mesh_e2n", "Mesh", "Point", "Solid", "collapse",
           "elastodynamic_", "elastoplasm", "elastoplastic_", "ic_collapse", "ic_ensemble", "ic_slump", "slump", "slump_",
           "hip", "grf"]
