"""MI355X-native sea-ice step with the Veris plugin interface.

``__VEROS_INTERFACE__`` mirrors ``veris/__init__.py:17-23`` (same name,
settings and variables; setup and run entry points), without requiring Veros.
The GPU drop-ins live in the modules named like the reference's
(``veris_amd.evp_solver.evp_solver``, ``veris_amd.growth.Growth``, ...) and in
:mod:`veris_amd.kernels`; :func:`veris_amd.step.seaice_step` runs a whole step.
"""

from .model import model
from .set_inits import set_inits
from .settings import SETTINGS
from .variables import VARIABLES

__version__ = "0.1.0"

__VEROS_INTERFACE__ = dict(
    name="veris",
    setup_entrypoint=set_inits,
    run_entrypoint=model,
    settings=SETTINGS,
    variables=VARIABLES,
)
