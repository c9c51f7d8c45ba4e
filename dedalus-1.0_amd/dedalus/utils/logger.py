"""The package log, ``mylog`` (the name the reference exports from
dedalus/utils/logger.py and its samples use).

One handler routes every record by level: debug lines go to stderr from
every process, info and above to stdout from process 0 only, each line
tagged with the process rank.  The level comes from ``decfg utils.loglevel``.
"""

import logging
import sys

from dedalus.config import decfg
from dedalus.utils.parallelism import com_sys

LINE = "%(asctime)s Dedalus [%(levelname)s] rank %(rank)d: %(message)s"


class RankRouter(logging.Handler):
    """Debug records to stderr on every rank; the rest to stdout on rank 0."""

    def __init__(self, rank):
        logging.Handler.__init__(self, logging.DEBUG)
        self.rank = rank
        self.setFormatter(logging.Formatter(LINE))

    def emit(self, record):
        if record.levelno > logging.DEBUG and self.rank != 0:
            return
        stream = sys.stderr if record.levelno <= logging.DEBUG else sys.stdout
        try:
            stream.write(self.format(record) + "\n")
            stream.flush()
        except Exception:
            self.handleError(record)


_base = logging.getLogger("Dedalus")
_base.setLevel(getattr(logging, decfg.get('utils', 'loglevel').upper()))
if not any(isinstance(h, RankRouter) for h in _base.handlers):
    _base.addHandler(RankRouter(com_sys.myproc))
_base.propagate = False

mylog = logging.LoggerAdapter(_base, {'rank': com_sys.myproc})
