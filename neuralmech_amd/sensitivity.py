"""``cuwave.sensitivity`` names of the call sites (``projects/8_physics_drivers/topopt_acoustic2D.py:15``,
``projects/4_specialized_nns/analog_rnn_fixture.py:9``): the boundary-reconstruction adjoint of the
variable mass / stiffness model."""

from .acoustic import reconstruction_nodes, reconstruction_sensitivity

__all__ = ["reconstruction_nodes", "reconstruction_sensitivity"]
