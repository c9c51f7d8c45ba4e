"""Auxiliary scalar ODE holder (dedalus/data_objects/aux_equation.py:25-39)."""


class AuxEquation:
    def __init__(self, RHS, kwargs, init_cond=0., order=1):
        self.order = order
        self.value = init_cond
        self._RHS = RHS
        self.kwargs = kwargs

    def RHS(self, value):
        return self._RHS(value, **self.kwargs)
