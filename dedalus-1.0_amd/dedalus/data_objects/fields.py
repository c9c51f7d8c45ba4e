"""Scalar, vector and tensor fields of representation components.

Same classes and semantics as dedalus/data_objects/fields.py:35-245; the
reductions and the solenoidal projection run on the device.
"""

import weakref

import numpy as np

from dedalus.data_objects.representations import (FourierRepresentation,
                                                  FourierShearRepresentation)
from dedalus.utils.logger import mylog
from dedalus.utils.parallelism import reduce_max


def create_field_classes(representation, shape, length):
    """Bind representation, shape and length to field classes (fields.py:35-57)."""
    attr_dict = {'representation': representation, 'shape': shape, 'length': length}
    return {"TensorField": type("TensorField", (TensorFieldBase,), dict(attr_dict)),
            "VectorField": type("VectorField", (VectorFieldBase,), dict(attr_dict)),
            "ScalarField": type("ScalarField", (ScalarFieldBase,), dict(attr_dict))}


class BaseField:
    """Base class for all field objects (fields.py:59-132)."""

    def __init__(self, sd, ncomp):
        self.sd = weakref.proxy(sd)
        self.ncomp = ncomp
        self.ndim = len(self.shape)
        self.components = [self.representation(self.sd, self.shape, self.length)
                           for _ in range(self.ncomp)]
        self.ctrans = {}

    def __getitem__(self, item):
        if isinstance(item, str):
            item = self.ctrans[item]
        return self.components[item]

    def __setitem__(self, item, data):
        if isinstance(item, str):
            item = self.ctrans[item]
        self.components[item] = data

    def __iter__(self):
        for i in range(self.ncomp):
            yield (i, self.components[i])

    def zero(self, comp, space='kspace'):
        if isinstance(comp, str):
            comp = self.ctrans[comp]
        self.components[comp][space] = 0.

    def zero_all(self, space='kspace'):
        for c in self.components:
            c[space] = 0.

    def save(self, group, prefix):
        group[prefix + '.representation'] = np.array(self.representation.__name__)
        group[prefix + '.type'] = np.array(self.__class__.__name__)
        for i, c in self:
            c.save(group, "%s/%i" % (prefix, i))

    def report_counts(self):
        for i, c in self:
            mylog.debug("component[%i] forward count = %i" % (i, c.fwd_count))
            mylog.debug("component[%i] rev count = %i" % (i, c.rev_count))


class TensorFieldBase(BaseField):
    """ndim**2 components (fields.py:134-139)."""

    def __init__(self, sd):
        BaseField.__init__(self, sd, sd.ndim ** 2)


class VectorFieldBase(BaseField):
    """ndim components (fields.py:141-195)."""

    def __init__(self, sd):
        BaseField.__init__(self, sd, sd.ndim)
        if self.ncomp == 2:
            self.ctrans = {'x': 0, 0: 'x', 'y': 1, 1: 'y'}
        else:
            self.ctrans = {'x': 0, 0: 'x', 'y': 1, 1: 'y', 'z': 2, 2: 'z'}

    def max_square(self):
        """max over components of max(c^2) in x space, reduced on the device."""
        c2 = np.array([c._plan.xreduce(c._xdev(), 0) for i, c in self])
        return reduce_max(c2, reduce_all=True)

    def l2norm(self):
        """sqrt(sum comp^2) in x space (fields.py:159-167)."""
        c0 = self.components[0]
        tmp = c0._plan.alloc_x(zero=False)
        c0._plan.xcombine(tmp, [(c._xdev(), c._xdev(), 1.0) for i, c in self][:4])
        return np.sqrt(c0._plan.download_x(tmp))

    def div_free(self):
        """Project off the irrotational part in k space (fields.py:169-195)."""
        if self.representation not in (FourierRepresentation, FourierShearRepresentation):
            raise NotImplementedError("Solenoidal projection not implemented for this representation.")
        mylog.debug("Performing solenoidal projection.")
        bufs = [c._kdev() for i, c in self]
        self.components[0]._plan.project(bufs)
        for i, c in self:
            c._mark('kspace')


class ScalarFieldBase(BaseField):
    """One component that acts as its representation (fields.py:197-245)."""

    def __init__(self, sd):
        BaseField.__init__(self, sd, 1)
        self.ctrans = {'': 0, 0: ''}

    def __getitem__(self, item):
        if item == 0:
            return self.components[0]
        return self.components[0][item]

    def __setitem__(self, item, data):
        if item == 0:
            self.components[0] = data
        else:
            self.components[0][item] = data

    def __getattr__(self, attr):
        if attr == 'components':
            raise AttributeError(attr)
        return getattr(self.components[0], attr)

    def zero(self, space='kspace'):
        self.components[0][space] = 0.
