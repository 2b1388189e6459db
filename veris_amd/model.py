"""Plugin run entry point (``veris/model.py:1-9``).

The reference's run entry point does nothing: the sea-ice step has to run
before the ocean step, so the host setup calls it from its forcing routine
(here :func:`veris_amd.step.seaice_step`).
"""


def model(state):
    return None
