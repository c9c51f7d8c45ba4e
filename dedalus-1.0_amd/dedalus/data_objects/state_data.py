"""State container: ordered fields plus time (dedalus/data_objects/state_data.py:45-186)."""

from collections import OrderedDict

import numpy as np

from dedalus.data_objects.fields import create_field_classes
from dedalus.utils.logger import mylog
from dedalus.utils.timer import Timer


class StateData:
    """All fields of a system state."""

    timer = Timer()

    def __init__(self, time, shape, length, field_class_dict, field_list=(), params=None):
        self.time = time
        self.shape = shape
        self.ndim = len(self.shape)
        self.length = length
        self._field_classes = field_class_dict
        self.parameters = params if params is not None else {}
        self.fields = OrderedDict()
        for name, ftype in field_list:
            self.add_field(name, ftype)

    def __getitem__(self, item):
        return self.fields[item]

    def __iter__(self):
        for name, field in self.fields.items():
            yield (name, field)

    def clone(self):
        return self.__class__(self.time, self.shape, self.length, self._field_classes,
                              params=self.parameters)

    def set_time(self, time):
        self.time = time

    def add_field(self, name, fieldtype):
        """Add a new field (allocates its device buffers)."""
        if name in self.fields:
            raise ValueError("Field with this name already exists.")
        self.fields[name] = self._field_classes[fieldtype](self)

    def snapshot(self, group):
        """Write every field into a dict of arrays (state_data.py:147-160)."""
        for name, field in self.fields.items():
            field.save(group, "fields/%s" % name)

    def create_tmp_data(self, space):
        fi = self.fields[next(iter(self.fields))][0]
        return np.zeros(tuple(fi.local_shape[space]), fi.dtype[space])

    def report_counts(self):
        for name, field in self.fields.items():
            mylog.debug("field %s" % name)
            field.report_counts()
