"""Logging adapter with the process rank (dedalus/utils/logger.py:49-72)."""

import logging
import sys

from dedalus.config import decfg
from dedalus.utils.parallelism import com_sys


class SingleLevelFilter(logging.Filter):
    def __init__(self, passlevel, reject):
        logging.Filter.__init__(self)
        self.passlevel = passlevel
        self.reject = reject

    def filter(self, record):
        if self.reject:
            return record.levelno != self.passlevel
        return record.levelno == self.passlevel


ufstring = "%(asctime)s %(name)-3s: [%(levelname)-9s] %(proc)i %(message)s"
formatter = logging.Formatter(ufstring)
extras = {'proc': com_sys.myproc}

dedaluslog = logging.getLogger("Dedalus")
dedaluslog.setLevel(getattr(logging, decfg.get('utils', 'loglevel').upper()))
mylog = logging.LoggerAdapter(dedaluslog, extras)

if not dedaluslog.handlers:
    debug_handler = logging.StreamHandler(sys.stderr)
    debug_handler.setLevel(logging.DEBUG)
    debug_handler.setFormatter(formatter)
    debug_handler.addFilter(SingleLevelFilter(logging.DEBUG, False))
    dedaluslog.addHandler(debug_handler)
    if com_sys.myproc == 0:
        info_handler = logging.StreamHandler(sys.stdout)
        info_handler.setLevel(logging.INFO)
        info_handler.setFormatter(formatter)
        dedaluslog.addHandler(info_handler)
