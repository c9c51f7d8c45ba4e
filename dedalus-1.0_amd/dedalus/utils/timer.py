"""Cumulative wall-clock timer decorator (dedalus/utils/timer.py:29-65)."""

import time
from functools import wraps

from dedalus.utils.parallelism import com_sys


class Timer:

    def __init__(self):
        self.timers = {}

    def __call__(self, func):
        @wraps(func)
        def wrapper(*args, **kwargs):
            start = time.time()
            retval = func(*args, **kwargs)
            stop = time.time()
            self.timers[func.__name__] = self.timers.get(func.__name__, 0.) + (stop - start)
            return retval
        return wrapper

    def print_stats(self, proc=0):
        if com_sys.myproc == proc:
            print()
            print("---Timings (proc %i)---" % proc)
            for k, v in self.timers.items():
                print("%s: %10.5f sec" % (k, v))
            print()


timer = Timer()
