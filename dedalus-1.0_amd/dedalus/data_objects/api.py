"""Public data-object API (dedalus/data_objects/api.py)."""

from .fields import create_field_classes
from .state_data import StateData
from .representations import (FourierRepresentation, FourierShearRepresentation,
                              ChebyshevRepresentation)
from .aux_equation import AuxEquation
